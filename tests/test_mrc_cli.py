"""MRC2014 reading (modes 0/1/2/6/12, both byte orders, extended-header
tilts) and float32 writing of the GCtfFind command line, exercised through
its --mrcinfo mode (no GPU needed)."""
import os
import subprocess

import numpy as np
import pytest

from mrc_util import read_mrc, write_mrc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gctffind_amd", "GCtfFind")


def run(*args):
    if not os.path.exists(CLI):
        pytest.skip("GCtfFind not built")
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=60)


def parse(out):
    lines = out.strip().splitlines()
    head = lines[0].split()
    h = {head[i]: int(head[i + 1]) for i in range(0, len(head), 2)}
    secs = []
    for ln in lines[1:]:
        t = ln.split()
        secs.append(dict(mean=float(t[3]), first=float(t[5]), last=float(t[7]), has_tilt=int(t[9]),
                         tilt=float(t[10])))
    return h, secs


@pytest.mark.parametrize("mode", [0, 1, 2, 6, 12])
@pytest.mark.parametrize("big", [False, True])
def test_read_modes(tmp_path, mode, big):
    rng = np.random.default_rng(mode)
    if mode == 0:
        d = rng.integers(0, 255, (3, 20, 30))
    elif mode == 1:
        d = rng.integers(-3000, 3000, (3, 20, 30))
    elif mode == 6:
        d = rng.integers(0, 60000, (3, 20, 30))
    else:
        d = rng.standard_normal((3, 20, 30)) * 10
    f = str(tmp_path / "a.mrc")
    write_mrc(f, d, mode=mode, big_endian=big, tilts=[-3.0, 0.0, 3.0])
    r = run("--mrcinfo", f)
    assert r.returncode == 0, r.stdout + r.stderr
    h, secs = parse(r.stdout)
    assert (h["nx"], h["ny"], h["nz"], h["mode"], h["swapped"]) == (30, 20, 3, mode, int(big))
    ref = d.astype(np.float16).astype(np.float64) if mode == 12 else d.astype(np.float32).astype(np.float64)
    for z, s in enumerate(secs):
        assert s["mean"] == pytest.approx(ref[z].mean(), rel=1e-6, abs=1e-6)
        assert s["first"] == pytest.approx(ref[z, 0, 0], rel=1e-6) and s["last"] == pytest.approx(ref[z, -1, -1], rel=1e-6)
        assert s["has_tilt"] == 1 and s["tilt"] == pytest.approx([-3.0, 0.0, 3.0][z])
    # every value: the sections converted to float32 (read in place into the
    # output buffer) and written back as a float32 stack
    out = str(tmp_path / "f32.mrc")
    r = run("--mrcinfo", f, out)
    assert r.returncode == 0, r.stdout + r.stderr
    assert np.array_equal(read_mrc(out), ref.astype(np.float32))


def test_float32_writer_round_trip(tmp_path):
    d = np.random.default_rng(1).standard_normal((2, 16, 24)).astype(np.float32)
    src, dst = str(tmp_path / "in.mrc"), str(tmp_path / "out.mrc")
    write_mrc(src, d, mode=2)
    r = run("--mrcinfo", src, dst)
    assert r.returncode == 0
    back = read_mrc(dst)
    assert np.array_equal(back, d)


def test_unsupported_and_missing(tmp_path):
    r = run("--mrcinfo", str(tmp_path / "nope.mrc"))
    assert r.returncode != 0
    r = run("--version")
    assert "GCtfFind" in r.stdout
    r = run("-InMrc", "a.mrc", "-OutCtf", "A.MRC")     # same names (CCtfFindMain.cpp:50-62)
    assert r.returncode == 1 and "same name" in r.stderr


def _patch_header(path, fmt, off, *vals):
    import struct
    b = bytearray(open(path, "rb").read())
    struct.pack_into(fmt, b, off, *vals)
    open(path, "wb").write(bytes(b))


@pytest.mark.parametrize("case", ["huge", "negative_ext", "truncated", "ext_past_end"])
def test_header_that_does_not_fit_the_file_is_refused(tmp_path, case):
    """A header whose sections (1024 + extended header + nz * nx * ny * bytes)
    do not fit in the file, or with a negative extended-header size, is
    refused when the file is opened: no section buffer is sized from it, and
    the command reports an error instead of allocating terabytes."""
    f = str(tmp_path / "a.mrc")
    d = np.random.default_rng(3).standard_normal((2, 20, 30)).astype(np.float32)
    write_mrc(f, d, mode=2)
    if case == "huge":
        _patch_header(f, "<3i", 0, 1 << 19, 1 << 19, 2)      # 2 TiB of float32 promised
    elif case == "negative_ext":
        _patch_header(f, "<i", 92, -1024)
    elif case == "truncated":
        open(f, "wb").write(open(f, "rb").read()[:1024 + 4 * 30 * 20 + 100])
    else:
        _patch_header(f, "<i", 92, 1 << 20)
    r = run("--mrcinfo", f)
    assert r.returncode == 1 and r.stdout.strip() == "error", r.stdout + r.stderr


def test_header_that_fits_exactly_is_read(tmp_path):
    """The size check accepts a file that holds exactly the promised sections."""
    f = str(tmp_path / "a.mrc")
    d = np.arange(2 * 4 * 6, dtype=np.float32).reshape(2, 4, 6)
    write_mrc(f, d, mode=2, tilts=[1.0, 2.0])
    r = run("--mrcinfo", f)
    assert r.returncode == 0, r.stdout + r.stderr
    h, secs = parse(r.stdout)
    assert (h["nx"], h["ny"], h["nz"]) == (6, 4, 2)
    assert secs[1]["last"] == 47.0
