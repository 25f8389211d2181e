"""gctf_do2d_batch_static (the search on device spectra that allocates
nothing) against gctf_set_half_spect + gctf_do2d, spectrum by spectrum,
through the C-ABI: results bit for bit equal, on two calls in a row (the
second with fewer spectra, so the static scratch still holds the first
call's contents), plus the size limit and the call-order errors."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

KEYS = ("df_min", "df_max", "azimuth_deg", "ext_phase_deg", "score", "ctf_res")


def _setup():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gctffind_amd import gctf
    L = gctf.load()
    L.gctf_do2d_batch_static.restype = ctypes.c_int
    L.gctf_do2d_batch_static.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    return gctf, L


def _spectra(gctf, n, pix):
    """Background-removed half spectra of n synthetic 1024^2 micrographs at
    pix (no rescale), from the single-micrograph spectrum path."""
    from gctffind_amd import synth
    out = []
    for i in range(n):
        img = synth.micrograph_np(1024, 1024, pix=pix, seed=700 + i, df1=9000.0 + 3100.0 * i + 300.0,
                                  df2=9000.0 + 3100.0 * i - 300.0, az_deg=-70.0 + 33.0 * i)
        ctx = gctf.Context()
        nnx, nny, _, pixn = ctx.rescale_setup(1024, 1024, pix)
        d = torch.from_numpy(img).cuda()
        ctx.rescale(d.data_ptr(), on_device=True)
        ctx.setup1(300.0, 2.7, 0.07, pixn)
        ctx.setup2(nnx, nny)
        ctx.gen_half_spectrum(ctx.scaled_image())
        out.append(ctx.get_half_spect(raw=False))
        ctx.close()
    return np.stack(out), pixn


def _reference(gctf, spects, pixn):
    """gctf_do2d on each spectrum (SetHalfSpect + Do2D after a fresh Setup1)."""
    ctx = gctf.Context()
    res = []
    for sp in spects:
        ctx.setup1(300.0, 2.7, 0.07, pixn)
        ctx.set_half_spect(sp)
        r = ctx.do2d()
        res.append(np.array([r["df_min"], r["df_max"], r["azimuth"], r["ext_phase"], r["score"], r["ctf_res"]],
                            np.float32))
    ctx.close()
    return np.stack(res)


def _static(L, ctx, spects):
    n = spects.shape[0]
    d = torch.from_numpy(np.ascontiguousarray(spects, np.float32)).cuda()
    out = torch.full((n, len(KEYS)), float("nan"), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ctx.check(L.gctf_do2d_batch_static(ctx.h, d.data_ptr(), n, out.data_ptr()))
    ctx.check(L.gctf_synchronize(ctx.h))
    return out.cpu().numpy()


def test_static_search_equals_do2d_two_calls():
    gctf, L = _setup()
    spects, pixn = _spectra(gctf, 6, 1.2)
    ref = _reference(gctf, spects, pixn)
    ctx = gctf.Context()
    ctx.setup1(300.0, 2.7, 0.07, pixn)
    first = _static(L, ctx, spects)
    print("first call (6 spectra)\n", first, flush=True)
    assert np.array_equal(first, ref), (first, ref)
    # a second, smaller call: the scratch still holds the first call's states
    pick = [4, 1, 5]
    second = _static(L, ctx, spects[pick])
    print("second call (3 spectra)\n", second, flush=True)
    assert np.array_equal(second, ref[pick]), (second, ref[pick])
    ctx.close()


def test_static_search_limits():
    gctf, L = _setup()
    ctx = gctf.Context()
    dummy = torch.zeros((17, 512, 257), dtype=torch.float32, device="cuda")
    out = torch.zeros((17, len(KEYS)), dtype=torch.float32, device="cuda")
    # before Setup1: a state error, nothing launched
    rc = L.gctf_do2d_batch_static(ctx.h, dummy.data_ptr(), 1, out.data_ptr())
    assert rc == gctf.GCTF_ERR_STATE
    ctx.setup1(300.0, 2.7, 0.07, 1.2)
    rc = L.gctf_do2d_batch_static(ctx.h, dummy.data_ptr(), 17, out.data_ptr())
    assert rc == gctf.GCTF_ERR_ARG
    assert b"at most 16" in L.gctf_error_string(ctx.h)
    assert L.gctf_do2d_batch_static(ctx.h, None, 1, out.data_ptr()) == gctf.GCTF_ERR_ARG
    assert L.gctf_do2d_batch_static(ctx.h, dummy.data_ptr(), 0, out.data_ptr()) == gctf.GCTF_ERR_ARG
    ctx.close()
    # tile 256: not supported by this entry point
    c2 = gctf.Context(tile=256)
    c2.setup1(300.0, 2.7, 0.07, 1.2)
    assert L.gctf_do2d_batch_static(c2.h, dummy.data_ptr(), 1, out.data_ptr()) == gctf.GCTF_ERR_UNSUPPORTED
    c2.close()
