"""Index contracts, the flatness gate's two branches, the phase-plate and
log-spectrum options and the full-size single-micrograph configuration,
HIP path against the CPU oracle and the committed fixtures.

  * radial average: every one of the 257 bins equals the mean over the
    committed FP32 sample table (tests/golden/radial_table_257x512.npz) on
    a spectrum of distinct pseudo-random values, within 5e-6 (the kernel
    sums in FP32 in table order, as GRadialAvg.cu does: up to 180 terms);
    one sample read from a different pixel moves its bin by about 2e-3
  * 2D correlation ring: the library's ring equals the committed index
    fixtures at 1.2 and 1.35 A (K = 47905 and 60793, the survey's counts)
  * flatness gate (CFindCtfBase.cpp:167-174): a spectrum with variance >= 1
    is low-passed, a flat one (a micrograph scaled by 1e-3) is not; both
    branches give the oracle's spectrum and search result
  * -ExtPhase with a range (CFindDefocus1D.cpp:84-103, CFindDefocus2D.cpp:
    324-361) and -LogSpect 1 (GRmBackground2D.cu:64-67): same results as the
    oracle, single image and batch
  * the 4096^2 float32 micrograph at 1.0 A (df 15500 / 14500 A, azimuth 30,
    SNR 0.1, B 100) with the 3000-40000 A / 100 A 1D scan, GPU batch path
    against the oracle's full estimate
Tolerances: defocus 5 A, azimuth 0.5 degrees, CC 1e-4, spectra 2e-5 (raw)
and 5e-5 (background-removed)."""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _setup():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gctffind_amd import gctf
    gctf.load()
    return gctf


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def _close(rg, ro, tag=""):
    assert abs(rg["df_min"] - ro["df_min"]) <= 5.0, (tag, rg, ro)
    assert abs(rg["df_max"] - ro["df_max"]) <= 5.0, (tag, rg, ro)
    assert abs(rg["azimuth"] - ro["azimuth"]) <= 0.5, (tag, rg, ro)
    assert abs(rg["ext_phase"] - ro["ext_phase"]) <= 0.5, (tag, rg, ro)
    assert abs(rg["score"] - ro["score"]) <= 1e-4, (tag, rg, ro)


def test_radial_average_every_bin_on_the_fixture_table():
    gctf = _setup()
    z = np.load(os.path.join(GOLDEN, "radial_table_257x512.npz"), allow_pickle=False)
    counts, xy = z["counts"].astype(np.int64), z["xy"].astype(np.int64)
    assert counts.size == 257 and counts.sum() == xy.shape[0]
    # distinct pseudo-random values: a sample from another pixel changes its
    # bin's mean by about 1 / count of the value range
    idx = np.arange(512 * 257, dtype=np.uint64)
    spect = (((idx * np.uint64(2654435761)) % np.uint64(1 << 20)).astype(np.float64) / (1 << 20) + 0.5)
    spect = spect.reshape(512, 257).astype(np.float32)
    want = np.empty(257)
    o = 0
    for r in range(257):
        c = counts[r]
        want[r] = spect[xy[o:o + c, 1], xy[o:o + c, 0]].astype(np.float64).mean()
        o += c
    ctx = gctf.Context()
    d = torch.from_numpy(np.stack([spect, spect[::-1].copy()])).cuda()
    out = torch.zeros((2, 257), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ctx.radial_average(d.data_ptr(), 2, out.data_ptr())
    got = out.cpu().numpy()
    assert np.abs(got[0] - want).max() <= 5e-6 * np.abs(want).max(), np.argmax(np.abs(got[0] - want))
    # the oracle's FP64 average over its own table: the same bins
    assert np.allclose(got[0], O.radial_avg(spect.astype(np.float64)), rtol=5e-6, atol=0)
    assert np.allclose(got[1], O.radial_avg(spect[::-1].astype(np.float64)), rtol=5e-6, atol=0)


@pytest.mark.parametrize("pix,count", [(1.2, 47905), (1.35, 60793)])
def test_ring_equals_fixture(pix, count):
    gctf = _setup()
    ref = np.load(os.path.join(GOLDEN, f"ring_pix{pix}_t512.npy"), allow_pickle=False)
    got = gctf.Context().ring_pixels(pix)
    assert ref.size == count and got.size == count
    assert np.array_equal(got, ref)


def _spectra(gctf, img, pix, **kw):
    ctx = gctf.Context(**kw)
    ny, nx = img.shape
    nnx, nny, pitch, pixn = ctx.rescale_setup(nx, ny, pix)
    d = torch.from_numpy(np.ascontiguousarray(img, np.float32)).cuda()
    ctx.rescale(d.data_ptr(), on_device=True)
    ctx.setup1(300.0, 2.7, 0.07, pixn)
    ctx.setup2(nnx, nny)
    ctx.gen_half_spectrum(ctx.scaled_image())
    return ctx, pixn


@pytest.mark.parametrize("scale,lowpassed", [(1.0, True), (1e-3, False)])
def test_flatness_gate_both_branches(scale, lowpassed):
    gctf = _setup()
    from gctffind_amd import synth
    pix = 1.2     # no rescale: the spectrum scales with the image
    img = synth.micrograph_np(1024, 1024, pix=pix, seed=91, df1=14000.0, df2=13400.0, az_deg=35.0)
    img = (img * np.float32(scale)).astype(np.float32)
    ctx, pixn = _spectra(gctf, img, pix)
    raw = ctx.get_half_spect(raw=True)
    fc = O.FindCtf2D(pix=pixn)
    fc.set_raw_spect(raw.astype(np.float64))
    assert fc.lowpassed() == lowpassed
    _, sp_o, _ = fc.spects()
    sp_g = ctx.get_half_spect(raw=False)
    assert _rel(sp_g, sp_o) < 5e-5
    ctx.set_phase(0.0, 0.0)
    fc.set_phase(0.0, 0.0)
    _close(ctx.do2d(), fc.do2d(), f"scale {scale}")


def test_phase_search_matches_oracle_on_same_spectrum():
    """-ExtPhase 45 60: the 1D brute force over 37 phases x the defocus scan
    and the 2D search's mRefinePhase line, on the GPU's spectrum."""
    gctf = _setup()
    from gctffind_amd import synth
    pix = 1.2
    img = synth.micrograph_np(1024, 1024, pix=pix, seed=93, df1=11200.0, df2=10600.0, az_deg=-25.0,
                              ext_phase_deg=60.0)
    ctx, pixn = _spectra(gctf, img, pix)
    sp = ctx.get_half_spect(raw=False)
    fc = O.FindCtf2D(pix=pixn)
    fc.set_half_spect(sp.astype(np.float64))
    for p0, rng in ((45.0, 60.0), (90.0, 120.0)):
        ctx.set_phase(p0, rng)
        fc.set_phase(p0, rng)
        r1g, r1o = ctx.do1d(), fc.do1d()
        assert r1g["df_min"] == r1o["df_min"] and r1g["ext_phase"] == r1o["ext_phase"], (r1g, r1o)
        ctx.set_phase(p0, rng)
        fc.set_phase(p0, rng)
        rg, ro = ctx.do2d(), fc.do2d()
        _close(rg, ro, f"phase {p0} {rng}")
        assert abs(rg["ext_phase"] - 60.0) <= 30.0, rg


def test_phase_search_batch_end_to_end():
    """The batched path with -ExtPhase 45 60 equals the single-image path and
    the oracle's full estimate."""
    gctf = _setup()
    from gctffind_amd import synth
    pix = 1.0
    imgs = np.stack([synth.micrograph_np(1024, 1024, pix=pix, seed=94 + i, df1=13000.0 + 4000 * i,
                                         df2=12500.0 + 4000 * i, az_deg=10.0 + 30 * i, ext_phase_deg=50.0)
                     for i in range(2)])
    ctx = gctf.Context(ext_phase=(45.0, 60.0))
    d = torch.from_numpy(imgs).cuda()
    rb = ctx.estimate_batch(d.data_ptr(), 2, 1024, 1024, pix)
    for i in range(2):
        c1, _ = _spectra(gctf, imgs[i], pix)
        c1.set_phase(45.0, 60.0)
        rs = c1.do2d()
        for k in ("df_min", "df_max", "azimuth", "ext_phase", "score", "ctf_res"):
            assert rb[i][k] == rs[k], (i, k, rb[i][k], rs[k])
        ro, _ = O.estimate(imgs[i], pix, phase=(45.0, 60.0))
        _close(rb[i], ro, f"image {i}")


def test_log_spect_matches_oracle():
    """-LogSpect 1 switches the background to the plain box mean
    (GRmBackground2D.cu:64-67); spectra and results equal the oracle's."""
    gctf = _setup()
    from gctffind_amd import synth
    pix = 1.2
    img = synth.micrograph_np(1024, 1024, pix=pix, seed=95, df1=17500.0, df2=16800.0, az_deg=55.0)
    ctx, pixn = _spectra(gctf, img, pix, log_spect=1)
    raw = ctx.get_half_spect(raw=True)
    fc = O.FindCtf2D(pix=pixn, log_spect=1)
    fc.set_raw_spect(raw.astype(np.float64))
    _, sp_o, _ = fc.spects()
    sp_g = ctx.get_half_spect(raw=False)
    assert _rel(sp_g, sp_o) < 5e-5
    # the log mode differs from the default one
    fc0 = O.FindCtf2D(pix=pixn)
    fc0.set_raw_spect(raw.astype(np.float64))
    assert _rel(fc0.spects()[1], sp_o) > 1e-3
    ctx.set_phase(0.0, 0.0)
    fc.set_phase(0.0, 0.0)
    _close(ctx.do2d(), fc.do2d(), "log")
    # batch path in log mode: the single-image result
    d = torch.from_numpy(img[None].copy()).cuda()
    rb = gctf.Context(log_spect=1).estimate_batch(d.data_ptr(), 1, 1024, 1024, pix)[0]
    ctx.set_phase(0.0, 0.0)
    rs = ctx.do2d()
    for k in ("df_min", "df_max", "azimuth", "score"):
        assert rb[k] == rs[k], (k, rb[k], rs[k])


def test_full_size_4096_single_micrograph_matches_oracle():
    """The single-micrograph workload at its own size: 4096^2 float32 at
    1.0 A (rescaled to 3412^2 at 1.2 A, 144 tiles), df 15500 / 14500 A at 30
    degrees, B 100, SNR 0.1, seed 0, 1D scan 3000-40000 A in 100 A steps
    (371 points), on both sides."""
    gctf = _setup()
    from gctffind_amd import synth
    pix, scan = 1.0, (3000.0, 40000.0, 100.0)
    img = synth.micrograph_np(4096, 4096, pix=pix, seed=0, df1=15500.0, df2=14500.0, az_deg=30.0,
                              bfac=100.0, snr=0.1)
    ctx = gctf.Context()
    ctx.set_search1d(*scan)
    d = torch.from_numpy(img[None].copy()).cuda()
    rg = ctx.estimate_batch(d.data_ptr(), 1, 4096, 4096, pix)[0]
    print("gpu", rg, flush=True)
    ro, fc = O.estimate(img, pix, scan1d=scan)
    print("oracle", ro, flush=True)
    _close(rg, ro, "config 0")
    assert rg["ctf_res"] == pytest.approx(ro["ctf_res"], rel=1e-6)
    # against the generating values (azimuth of the larger defocus)
    assert abs(0.5 * (rg["df_min"] + rg["df_max"]) - 15000.0) <= 100.0
    assert abs(rg["azimuth"] - 30.0) <= 5.0
