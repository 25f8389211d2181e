"""The C-ABI library loads and exports every entry point include/gctf_hip.h
declares (no GPU work is started here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gctf_hip.h")
LIB = os.path.join(ROOT, "gctffind_amd", "libgctf_hip.so")


def declared():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"GCTF_API\s+[\w\s\*]+?\b(gctf_\w+)\s*\(", src)))


def test_header_declares_the_reference_surface():
    names = declared()
    for n in ("gctf_rescale_setup", "gctf_rescale", "gctf_scaled_image", "gctf_setup1",
              "gctf_setup2", "gctf_set_phase", "gctf_gen_half_spectrum", "gctf_do1d",
              "gctf_refine1d", "gctf_do2d", "gctf_refine", "gctf_spect_size",
              "gctf_get_half_spect", "gctf_set_half_spect", "gctf_full_spectrum",
              "gctf_estimate_batch"):
        assert n in names
    # every declaration cites a reference location or says it has none
    src = open(HEADER).read()
    assert src.count("CFindCTFInc.h") >= 5


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.skip("libgctf_hip.so not built")
    L = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing
    from gctffind_amd import gctf
    assert sorted(gctf.EXPORTS) == declared()


def test_create_without_gpu_fails_cleanly():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from gctffind_amd import gctf
    L = gctf.load()
    h = ctypes.c_void_p()
    rc = L.gctf_create(None, 0, ctypes.byref(h))
    assert rc != 0 and not h.value


def test_tile_grid_equals_oracle_and_shape_table():
    """The library's tile grid (host integer code, no GPU work) is the
    oracle's on every size of a sweep and the survey's shape table
    (CGenAvgSpectrum.cpp:53-63)."""
    if not os.path.exists(LIB):
        pytest.skip("libgctf_hip.so not built")
    import oracle as O
    from gctffind_amd import gctf
    table = {(3412, 3412): (12, 12, 42, 42), (3838, 3710): (13, 13, 127, 63), (4096, 4096): (15, 15, 0, 0),
             (2400, 1704): (8, 5, 48, 84), (1024, 1024): (3, 3, 0, 0)}
    for (nx, ny), want in table.items():
        assert gctf.tile_grid(nx, ny, 512)[:4] == want
    for tile in (256, 512, 1024):
        for nx in range(tile, 12000, 173):
            ny = tile + (nx * 7) % 9000
            assert gctf.tile_grid(nx, ny, tile) == O.tile_grid(nx, ny, tile), (nx, ny, tile)
