"""HIP path against the CPU oracle, stage by stage and end to end.

Every test feeds the same float32 input to libgctf_hip.so (through the C-ABI)
and to the oracle, and checks the documented tolerances:
  * index tables (tile grid, radial samples, ring): bit-exact
  * spectra: relative error <= 2e-5 (float32 FFTs against an FP64 oracle)
  * search results: defocus within 5 A, azimuth within 0.5 deg, CC within 1e-4
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gctffind_amd import gctf
    gctf.load()
    return torch.device("cuda:0")


def _ctx(**kw):
    from gctffind_amd import gctf
    return gctf.Context(**kw)


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def _img(n=1024, pix=1.2, seed=1, **kw):
    from gctffind_amd import synth
    return synth.micrograph_np(n, n, pix=pix, seed=seed, **kw)


@pytest.mark.parametrize("shape,pix", [((1024, 1024), 1.0), ((1000, 900), 0.8), ((768, 1024), 1.2),
                                       ((720, 960), 0.6), ((1200, 1500), 0.8),
                                       ((4096, 256), 1.0), ((256, 4096), 1.0), ((4092, 256), 0.8)])
def test_rescale_matches_oracle(shape, pix):
    """Includes lengths 2^a 3^b 5^c in every stage, which run as mixed-radix
    Stockham transforms: 960 x 720 at 0.6 A crops to 480 x 360, and
    1500 x 1200 at 0.8 A to 1000 x 800.  And lengths n = 4 q whose quarter
    has a quarter of n's Bluestein length, which run as one radix-4 step
    over four q-point Bluestein transforms: the 3412-point inverse columns
    (4096 high at 1.0 A) and C2R rows (4096 wide at 1.0 A) of the 4096^2
    workload, and 4092-point forward and 2728-point inverse columns."""
    ny, nx = shape
    _check_rescale(_img(max(shape), pix)[:ny, :nx].copy(), pix)


@pytest.mark.parametrize("shape,pix", [((256, 11520), 0.5), ((256, 11520), 1.0), ((256, 11520), 1.1),
                                       ((11520, 256), 0.5), ((11520, 256), 0.7), ((8184, 256), 0.5),
                                       ((8184, 256), 1.0), ((256, 5760), 0.8), ((256, 5760), 1.0)])
def test_rescale_long_rows_matches_oracle(shape, pix):
    """Rows of 11520 pixels (the super-resolution width): the R2C row transform
    runs at half length, even and odd samples packed into one complex row, as
    a 5760-point mixed-radix transform; at 1.0 A the C2R rows (9600 pixels)
    do as well.  Columns of 11520 pixels (portrait): the column pass splits
    its forward transform into those of the even and odd samples when the
    rescaled height fits 8192-point transforms (0.5 A), else runs it whole on
    the 16384-point kernel (0.7 A); columns of 8184 pixels (the
    super-resolution height) split onto 8192-point Bluestein transforms.
    Rows of 5760 pixels run as whole 5760-point mixed-radix transforms.  At
    1.1 A the 11520-pixel rows rescale to 10560, whose half-length C2R rows of
    5280 = 4 x 1320 run as four 1320-point Bluestein transforms on 4096 inside
    the 16384-point kernel."""
    from gctffind_amd import synth
    _check_rescale(synth.micrograph_np(shape[0], shape[1], pix=pix, seed=7), pix)


def test_rescale_split_columns_limit():
    """A split column transform whose rescaled height exceeds 8192 is refused
    with an error, not computed.  11522 = 2 x 7 x 823 has no mixed-radix
    plan: its columns split into two 5761-point Bluestein transforms, and at
    1.0 A the rescaled height is 9602.  (11520 = 2^8 3^2 5 instead fits one
    11520-point mixed-radix transform and needs no split.)"""
    _dev()
    from gctffind_amd import gctf
    with pytest.raises(RuntimeError, match="8192"):
        _ctx().rescale_setup(256, 11522, 1.0)
    nnx, nny, pitch, pixn = _ctx().rescale_setup(256, 11520, 1.0)
    assert nny == 9600


@pytest.mark.parametrize("nx", [11520, 5760])
def test_long_rows_batch_equals_single(nx):
    """Row partials of the half-length C2R rows (9600 and 4800 pixels) equal
    k_row_sums on the rescaled image: the batch (partials from the rescale)
    and the single path (partials from memory) give the same bits."""
    dev = _dev()
    from gctffind_amd import synth
    ny, pix = 640, 1.0
    img = synth.micrograph_np(ny, nx, pix=pix, seed=8, df1=16000.0, df2=15400.0, az_deg=40.0)
    d = torch.from_numpy(img).to(dev)
    rb = _ctx().estimate_batch(d.data_ptr(), 1, nx, ny, pix)[0]
    ctx1, _, _ = _gpu_spectra(img, pix)
    ctx1.set_phase(0.0, 0.0)
    rs = ctx1.do2d()
    for k in ("df_min", "df_max", "azimuth", "score", "ctf_res"):
        assert rb[k] == rs[k], (k, rb[k], rs[k])


def _check_rescale(img, pix):
    dev = _dev()
    ny, nx = img.shape
    ctx = _ctx()
    nnx, nny, pitch, pixn = ctx.rescale_setup(nx, ny, pix)
    onnx, onny, opitch, opixn, _ = O.rescale_sizes(nx, ny, pix)
    assert (nnx, nny, pitch) == (onnx, onny, opitch)
    assert pixn == pytest.approx(opixn, rel=0, abs=0)
    d = torch.from_numpy(img).to(dev)
    ctx.rescale(d.data_ptr(), on_device=True)
    ptr = ctx.scaled_image()
    torch.cuda.synchronize()
    host = np.empty((nny, pitch), np.float32)
    _copy_dev_to_host(ptr, host)
    ref = O.rescale(img, pix)
    assert _rel(host[:, :nnx], ref[:, :nnx]) < 2e-5
    # host-memory input takes the same path
    ctx.rescale(img, on_device=False)
    host2 = np.empty_like(host)
    _copy_dev_to_host(ctx.scaled_image(), host2)
    assert np.array_equal(host2[:, :nnx], host[:, :nnx])


@pytest.mark.parametrize("pix", [0.8, 1.2])
def test_rescale_prefetch_uses_the_copy_and_drops_a_stale_one(pix):
    """gctf_rescale_prefetch: the next rescale of the prefetched host image
    gives the same bits as an ordinary one; a rescale of another image drops
    the prefetch and copies its own (at 0.8 A through the rescale, at 1.2 A
    through the unrescaled row copy)."""
    _dev()
    a = _img(1024, pix, seed=11)[:, :1000].copy()
    b = _img(1024, pix, seed=12)[:, :1000].copy()
    ny, nx = a.shape
    ctx = _ctx()
    nnx, nny, pitch, _ = ctx.rescale_setup(nx, ny, pix)

    def scaled():
        h = np.empty((nny, pitch), np.float32)
        _copy_dev_to_host(ctx.scaled_image(), h)
        return h[:, :nnx].copy()

    ctx.rescale(a)
    ra = scaled()
    ctx.rescale(b)
    rb = scaled()
    assert not np.array_equal(ra, rb)
    for _ in range(2):
        ctx.rescale_prefetch(a)
        ctx.rescale(a)
        assert np.array_equal(scaled(), ra)
        ctx.rescale_prefetch(b)
        ctx.rescale(b)
        assert np.array_equal(scaled(), rb)
    # a prefetch of b, then a rescale of a: a's own copy, and b's is dropped
    ctx.rescale_prefetch(b)
    ctx.rescale(a)
    assert np.array_equal(scaled(), ra)
    ctx.rescale(b)
    assert np.array_equal(scaled(), rb)


def _copy_dev_to_host(ptr, host):
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.restype = ctypes.c_int
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    rc = hip.hipMemcpy(host.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr), host.nbytes, 2)
    assert rc == 0


def _gpu_spectra(img, pix, tile=512, log_spect=0):
    """GenHalfSpectrum on the GPU for a float32 image (rescaled first)."""
    dev = _dev()
    ny, nx = img.shape
    ctx = _ctx(tile=tile, log_spect=log_spect)
    nnx, nny, pitch, pixn = ctx.rescale_setup(nx, ny, pix)
    d = torch.from_numpy(np.ascontiguousarray(img, np.float32)).to(dev)
    ctx.rescale(d.data_ptr(), on_device=True)
    ctx.setup1(300.0, 2.7, 0.07, pixn)
    ctx.setup2(nnx, nny)
    ctx.gen_half_spectrum(ctx.scaled_image())
    pad = np.empty((nny, pitch), np.float32)
    _copy_dev_to_host(ctx.scaled_image(), pad)
    return ctx, pad, (nnx, nny, pixn)


@pytest.mark.parametrize("n,pix", [(1024, 1.2), (1024, 1.0), (1400, 1.35)])
def test_avg_spectrum_and_background(n, pix):
    img = _img(n, pix, seed=3)
    ctx, pad, (nnx, nny, pixn) = _gpu_spectra(img, pix)
    raw_g = ctx.get_half_spect(raw=True)
    raw_o = O.avg_spectrum(pad.astype(np.float64), nnx, nny)
    assert _rel(raw_g, raw_o) < 2e-5
    # background + gate + low-pass from the same raw spectrum
    fc = O.FindCtf2D(pix=pixn)
    fc.set_raw_spect(raw_g.astype(np.float64))
    _, sp_o, _ = fc.spects()
    sp_g = ctx.get_half_spect(raw=False)
    assert _rel(sp_g, sp_o) < 5e-5


def test_search_1d_2d_matches_oracle_on_same_spectrum():
    img = _img(1024, 1.2, seed=5, df1=21000.0, df2=20000.0, az_deg=-40.0)
    ctx, pad, (nnx, nny, pixn) = _gpu_spectra(img, 1.2)
    sp_g = ctx.get_half_spect(raw=False)
    fc = O.FindCtf2D(pix=pixn)
    fc.set_half_spect(sp_g.astype(np.float64))
    fc.set_phase(0.0, 0.0)
    ro = fc.do2d()
    ctx.set_phase(0.0, 0.0)
    rg = ctx.do2d()
    assert abs(rg["df_min"] - ro["df_min"]) <= 5.0
    assert abs(rg["df_max"] - ro["df_max"]) <= 5.0
    assert abs(rg["azimuth"] - ro["azimuth"]) <= 0.5
    assert abs(rg["score"] - ro["score"]) <= 1e-4
    # spectral-CC resolution: the same 0.143 shell (GSpectralCC2D.cu:157-171)
    assert rg["ctf_res"] == pytest.approx(ro["ctf_res"], rel=1e-6)


def test_search1d_configured_scan_matches_oracle():
    """Do1D on the 3000-40000 A / 100 A scan (371 points) lands where the
    oracle's does, on the same spectrum."""
    img = _img(1024, 1.2, seed=13, df1=16500.0, df2=15800.0, az_deg=20.0)
    ctx, pad, (nnx, nny, pixn) = _gpu_spectra(img, 1.2)
    sp_g = ctx.get_half_spect(raw=False)
    fc = O.FindCtf2D(pix=pixn)
    fc.set_half_spect(sp_g.astype(np.float64))
    fc.set_phase(0.0, 0.0)
    fc.set_search1d(3000.0, 40000.0, 100.0)
    ro = fc.do1d()
    ctx.set_phase(0.0, 0.0)
    ctx.set_search1d(3000.0, 40000.0, 100.0)
    rg = ctx.do1d()
    assert rg["df_min"] == ro["df_min"]
    assert abs(rg["score"] - ro["score"]) <= 1e-6
    ctx.set_search1d()
    fc.set_search1d()
    assert ctx.do1d()["df_min"] == fc.do1d()["df_min"]


@pytest.mark.parametrize("seed,df1,df2,az", [(31, 9000.0, 8600.0, 5.0), (32, 26000.0, 25000.0, -75.0),
                                              (33, 15000.0, 14900.0, 33.0), (35, 33000.0, 32000.0, 0.0)])
def test_search_result_identical_to_oracle(seed, df1, df2, az):
    """On the same spectrum the float32 search lands on the oracle's candidate:
    defocus and azimuth equal, the CC within one float32 step (the FP32
    per-thread sums against the oracle's FP64 ones)."""
    img = _img(1024, 1.2, seed=seed, df1=df1, df2=df2, az_deg=az)
    ctx, pad, (nnx, nny, pixn) = _gpu_spectra(img, 1.2)
    sp_g = ctx.get_half_spect(raw=False)
    fc = O.FindCtf2D(pix=pixn)
    fc.set_half_spect(sp_g.astype(np.float64))
    fc.set_phase(0.0, 0.0)
    ro = fc.do2d()
    ctx.set_phase(0.0, 0.0)
    rg = ctx.do2d()
    assert rg["df_min"] == ro["df_min"] and rg["df_max"] == ro["df_max"]
    assert rg["azimuth"] == ro["azimuth"]
    assert abs(rg["score"] - ro["score"]) <= 1e-7


def test_refine_matches_oracle():
    img = _img(1024, 1.2, seed=7, df1=12500.0, df2=12000.0, az_deg=60.0)
    ctx, pad, (nnx, nny, pixn) = _gpu_spectra(img, 1.2)
    sp_g = ctx.get_half_spect(raw=False)
    fc = O.FindCtf2D(pix=pixn)
    fc.set_half_spect(sp_g.astype(np.float64))
    args = ([12300.0, 5000.0], [0.01, 0.04], [50.0, 30.0], [0.0, 0.0])
    ro = fc.refine(*args)
    rg = ctx.refine(*args)
    for k in ("df_min", "df_max"):
        assert abs(rg[k] - ro[k]) <= 5.0
    assert abs(rg["azimuth"] - ro["azimuth"]) <= 0.5
    assert abs(rg["score"] - ro["score"]) <= 1e-4


@pytest.mark.parametrize("n,pix,seed", [(1024, 1.2, 11), (1024, 1.0, 12)])
def test_end_to_end_matches_oracle(n, pix, seed):
    img = _img(n, pix, seed=seed, df1=18000.0, df2=17200.0, az_deg=15.0)
    ctx, pad, (nnx, nny, pixn) = _gpu_spectra(img, pix)
    ctx.set_phase(0.0, 0.0)
    rg = ctx.do2d()
    ro, _ = O.estimate(img, pix)
    assert abs(rg["df_min"] - ro["df_min"]) <= 5.0
    assert abs(rg["df_max"] - ro["df_max"]) <= 5.0
    assert abs(rg["azimuth"] - ro["azimuth"]) <= 0.5
    assert abs(rg["score"] - ro["score"]) <= 1e-4
    assert rg["ctf_res"] == pytest.approx(ro["ctf_res"], rel=1e-6)


def test_batch_equals_single():
    dev = _dev()
    n, pix = 1024, 1.0
    imgs = np.stack([_img(n, pix, seed=20 + i, df1=14000.0 + 3000 * i, df2=13500.0 + 3000 * i,
                          az_deg=10.0 * i) for i in range(3)])
    d = torch.from_numpy(imgs).to(dev)
    ctx = _ctx()
    rb = ctx.estimate_batch(d.data_ptr(), 3, n, n, pix)
    rb2 = ctx.estimate_batch(d.data_ptr(), 3, n, n, pix)  # graph replay
    for i in range(3):
        ctx1, _, _ = _gpu_spectra(imgs[i], pix)
        ctx1.set_phase(0.0, 0.0)
        rs = ctx1.do2d()
        for k in ("df_min", "df_max", "azimuth", "score", "ctf_res"):
            assert rb[i][k] == rs[k], (i, k)
            assert rb2[i][k] == rs[k], (i, k)


def test_batch_uses_configured_scan():
    """The batched plain path runs the 1D scan set with set_search1d: a batch
    of one equals the single-micrograph Do2D with the same scan."""
    dev = _dev()
    n, pix = 1024, 1.0
    img = _img(n, pix, seed=41, df1=9000.0, df2=8700.0, az_deg=10.0)
    d = torch.from_numpy(img[None].copy()).to(dev)
    ctx = _ctx()
    ctx.set_search1d(3000.0, 40000.0, 100.0)
    rb = ctx.estimate_batch(d.data_ptr(), 1, n, n, pix)[0]
    ctx1, _, _ = _gpu_spectra(img, pix)
    ctx1.set_phase(0.0, 0.0)
    ctx1.set_search1d(3000.0, 40000.0, 100.0)
    rs = ctx1.do2d()
    for k in ("df_min", "df_max", "azimuth", "score", "ctf_res"):
        assert rb[k] == rs[k], (k, rb[k], rs[k])


def test_full_spectrum_matches_oracle():
    img = _img(1024, 1.2, seed=9, df1=16000.0, df2=15000.0, az_deg=70.0)
    ctx, pad, (nnx, nny, pixn) = _gpu_spectra(img, 1.2)
    sp_g = ctx.get_half_spect(raw=False)
    fc = O.FindCtf2D(pix=pixn)
    fc.set_half_spect(sp_g.astype(np.float64))
    fc.set_phase(0.0, 0.0)
    ro = fc.do2d()
    ctx.set_phase(0.0, 0.0)
    rg = ctx.do2d()
    assert rg["df_min"] == ro["df_min"] and rg["azimuth"] == ro["azimuth"]
    fg = ctx.full_spectrum()
    fo = fc.full_spectrum()
    assert fg.shape == fo.shape == (512, 512)
    scale = np.abs(fo).max()
    assert np.abs(fg - fo).max() <= 1e-4 * scale
    # left half carries the fitted |CTF|, right half the spectrum itself
    assert np.array_equal(fg[:, 256:], sp_g[(np.arange(512)), :256].astype(np.float32)) or \
        np.allclose(fg[:, 256:], fo[:, 256:], rtol=0, atol=1e-6 * scale)


def test_gpu_radial_and_ring_tables_match_oracle():
    # a random spectrum: any differing radial sample or ring pixel would move
    # the 1D/2D correlations far beyond the tolerance used here
    _dev()
    rng = np.random.default_rng(21)
    spect = rng.random((512, 257)).astype(np.float32)
    ctx = _ctx()
    ctx.setup1(300.0, 2.7, 0.07, 1.2)
    ctx.set_half_spect(spect)
    ctx.set_phase(0.0, 0.0)
    fc = O.FindCtf2D(pix=1.2)
    fc.set_half_spect(spect.astype(np.float64))
    rg = ctx.do1d()
    ro = fc.do1d()
    assert abs(rg["score"] - ro["score"]) <= 1e-6
    assert rg["df_min"] == ro["df_min"]
