"""Pins the CPU oracle (oracle/gctf_oracle.c) without the reference.

The reference ships no tests or golden vectors, and it cannot be built here
(nvcc, cuFFT and Lib/libcuutilfft.a are absent).  Each oracle function is
checked against something independent of its own code: numpy's FFT, brute
force sums written separately here, closed forms (the CTF zeros from the
CCTFTheory quadratic, Parseval, the Gaussian low-pass transfer), invariants
(scale invariance of the CC, constant-input behaviour) and the shape table
and FP32/FP64 radial-sample facts recorded in the survey of the reference.
"""
import math

import numpy as np
import pytest

import oracle as O


def rng(seed=0):
    return np.random.default_rng(seed)


# ---------------------------------------------------------------- a1 CTF params
def test_wavelength_closed_form():
    # 12.26 / sqrt(V (1 + 0.9784e-6 V)) A with V in volts (CCTFTheory.cpp:285)
    for kv in (100.0, 200.0, 300.0):
        v = kv * 1000.0
        assert O.wavelength(kv) == pytest.approx(12.26 / math.sqrt(v + 0.9784e-6 * v * v), rel=1e-7)
    assert O.wavelength(300.0) == pytest.approx(0.0196808, abs=5e-8)


def test_ctf_params_units():
    p = O.ctf_params(300.0, 2.7, 0.07, 1.2, 0.0)
    assert p[0] == pytest.approx(12.26 / math.sqrt(300e3 + 0.9784e-6 * 300e3 ** 2) / 1.2, rel=1e-6)
    assert p[1] == pytest.approx(2.7e7 / 1.2, rel=1e-7)         # Cs in pixels
    assert p[2] == pytest.approx(math.atan(0.07 / math.sqrt(1 - 0.07 ** 2)), rel=1e-6)
    # extra phase is reduced modulo pi (CCTFTheory.cpp:85)
    assert O.ctf_params(300, 2.7, 0.07, 1.0, 4.0)[3] == pytest.approx(4.0 - math.pi, abs=1e-6)


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [1, 2, 8, 30, 97, 512, 853, 1706, 3412, 4096])
def test_fft_matches_numpy(n):
    x = rng(n).standard_normal(n) + 1j * rng(n + 1).standard_normal(n)
    assert np.allclose(O.fft(x), np.fft.fft(x), rtol=0, atol=1e-10 * max(1, n))
    assert np.allclose(O.fft(x, inverse=True), np.fft.ifft(x) * n, rtol=0, atol=1e-10 * max(1, n))


# ---------------------------------------------------------------- a2 rescale
def _numpy_rescale(img, nnx, nny):
    """Fourier crop written with numpy: rfft2, keep kx < nnx/2+1 and
    ky <= nny/2 / the top nny/2-1 negative rows, then the real part of the
    inverse of the Hermitian extension (unnormalised both ways)."""
    ny, nx = img.shape
    F = np.fft.rfft2(img.astype(np.float64))
    ncx = nnx // 2 + 1
    rows = list(range(0, nny // 2 + 1)) + list(range(ny - nny + nny // 2 + 1, ny))
    G = F[rows][:, :ncx]
    full = np.zeros((nny, nnx), complex)
    full[:, :ncx] = G
    for k in range(ncx, nnx):
        full[:, k] = np.conj(G[(-np.arange(nny)) % nny, nnx - k])
    return np.real(np.fft.ifft2(full)) * nnx * nny


@pytest.mark.parametrize("shape,pix", [((64, 64), 1.0), ((100, 90), 0.8), ((70, 96), 0.55)])
def test_rescale_matches_numpy(shape, pix):
    img = rng(3).standard_normal(shape).astype(np.float32) + 5.0
    nnx, nny, pitch, pixn, b = O.rescale_sizes(shape[1], shape[0], pix)
    out = O.rescale(img, pix)
    assert out.shape == (nny, pitch)
    ref = _numpy_rescale(img, nnx, nny)
    assert np.allclose(out[:, :nnx], ref, rtol=0, atol=1e-9 * np.abs(ref).max())
    assert np.all(out[:, nnx:] == 0)


def test_rescale_brute_force_dft():
    img = rng(4).standard_normal((10, 12)).astype(np.float32)
    nnx, nny, pitch, pixn, b = O.rescale_sizes(12, 10, 0.8)
    assert (nnx, nny) == (8, 6)
    y, x = np.mgrid[0:10, 0:12]
    F = np.zeros((10, 7), complex)
    for ky in range(10):
        for kx in range(7):
            F[ky, kx] = np.sum(img * np.exp(-2j * np.pi * (kx * x / 12 + ky * y / 10)))
    keep = [0, 1, 2, 3, 8, 9]                     # ky <= 3, then the last two rows
    G = F[keep][:, :5]
    yy, xx = np.mgrid[0:6, 0:8]
    ref = np.zeros((6, 8))
    for ky in range(6):
        for kx in range(5):
            w = 1.0 if kx in (0, 4) else 2.0      # Hermitian pairs counted twice
            ref += w * np.real(G[ky, kx] * np.exp(2j * np.pi * (kx * xx / 8 + ky * yy / 6)))
    assert np.allclose(O.rescale(img, 0.8)[:, :8], ref, atol=1e-9)


def test_rescale_sizes_and_identity():
    # shape table of the survey (CRescaleImage.cpp:54-64 in float)
    assert O.rescale_sizes(4096, 4096, 1.0)[:3] == (3412, 3412, 3414)
    assert O.rescale_sizes(5760, 4092, 0.5)[:2] == (2400, 1704)
    assert O.rescale_sizes(5760, 4092, 0.8)[:2] == (3840, 2728)
    assert O.rescale_sizes(5760, 4092, 1.1)[:2] == (5280, 3750)
    nnx, nny, pitch, pixn, b = O.rescale_sizes(3838, 3710, 1.35)
    assert (nnx, nny, b) == (3838, 3710, 1.0) and pixn == pytest.approx(1.35)
    img = rng(5).standard_normal((33, 41)).astype(np.float32)
    out = O.rescale(img, 1.5)                      # no binning: rows copied
    assert np.array_equal(out[:32, :41], img[:32].astype(np.float64))


# ---------------------------------------------------------------- a4 tile grid
@pytest.mark.parametrize("size,expect", [
    ((3412, 3412), (12, 12, 42, 42)), ((3838, 3710), (13, 13, 127, 63)),
    ((4096, 4096), (15, 15, 0, 0)), ((2400, 1704), (8, 5, 48, 84)), ((1024, 1024), (3, 3, 0, 0))])
def test_tile_grid(size, expect):
    g = O.tile_grid(size[0], size[1], 512)
    assert g[:4] == expect and g[4] == 256 and g[5] == 256


# ---------------------------------------------------------------- a5-a10 spectrum
def _mask(t):
    y, x = np.mgrid[0:t, 0:t].astype(np.float64)
    c = t * 0.5
    r = np.hypot(2 * np.abs(x - c) / t, 2 * np.abs(y - c) / t)
    q = 0.5 * (1 - np.cos(float(np.float32(3.1415926)) * r))
    return np.where(r >= 1, 0.0, 1 - q ** 3)


def test_round_edge_closed_form():
    m = O.round_edge_table(64)
    assert np.allclose(m, _mask(64), atol=1e-15)
    assert m[32, 32] == 1.0 and np.all(m[0, :] == 0) and np.all(m[:, 0] == 0)
    assert np.allclose(m[1:, 1:], m[1:, 1:][::-1, ::-1])   # symmetric about the centre


def _numpy_avg_spectrum(pad, nx, ny, t):
    ntx, nty, ox, oy, ov, step = O.tile_grid(nx, ny, t)
    m = _mask(t)
    acc = np.zeros((t, t // 2 + 1))
    fac = float(np.float32(1.0) / np.float32(ntx * nty))
    for j in range(nty):
        for i in range(ntx):
            tile = pad[oy + j * step: oy + j * step + t, ox + i * step: ox + i * step + t]
            a = np.abs(np.fft.rfft2((tile - tile.mean()) * m))
            acc += np.roll(a, t // 2, axis=0) * fac
    return acc


@pytest.mark.parametrize("t,nx,ny", [(64, 200, 150), (16, 40, 40), (512, 1024, 768)])
def test_avg_spectrum_matches_numpy(t, nx, ny):
    img = rng(6).standard_normal((ny, nx)) * 3 + 7
    pad = np.zeros((ny, (nx // 2 + 1) * 2))
    pad[:, :nx] = img
    out = O.avg_spectrum(pad, nx, ny, t)
    assert np.allclose(out, _numpy_avg_spectrum(pad, nx, ny, t), rtol=1e-10, atol=1e-9)


def test_tile_spectrum_parseval_single_tile():
    t = 32
    img = rng(7).standard_normal((t, t))
    pad = np.zeros((t, t + 2))
    pad[:, :t] = img
    s = O.avg_spectrum(pad, t, t, t)               # one tile, DC moved to row t/2
    full = np.roll(s, -(t // 2), axis=0)
    e_half = np.sum(full[:, 0] ** 2) + np.sum(full[:, t // 2] ** 2) + 2 * np.sum(full[:, 1:t // 2] ** 2)
    w = (img - img.mean()) * _mask(t)
    assert e_half == pytest.approx(t * t * np.sum(w ** 2), rel=1e-10)


# ---------------------------------------------------------------- a11 background
def _numpy_background(s, hb, log=False):
    cy, cx = s.shape
    t = cy
    P = np.empty((t, t))                           # periodic Friedel-unfolded plane
    for Y in range(t):
        for X in range(t):
            P[Y, X] = s[Y, X] if X < t // 2 else s[(t - Y) % t, t - X]
    P = P if log else np.sqrt(P)
    B = np.zeros_like(P)
    for k in range(-hb, hb + 1):
        for j in range(-hb, hb + 1):
            B += np.roll(np.roll(P, -k, axis=0), -j, axis=1)
    B /= (2 * hb + 1) ** 2
    if not log:
        B = B * B
    out = s - B[:, :cx]
    y, x = np.mgrid[0:cy, 0:cx]
    r = np.sqrt((x.astype(np.float32) * np.float32(0.5) / np.float32(cx - 1)) ** 2
                + ((y - cy // 2) ** 2).astype(np.float32) / np.float32(cy * cy)).astype(np.float32)
    out[r.astype(np.float64) < 0.04] = 0.0
    return out


@pytest.mark.parametrize("log", [0, 1])
def test_background_matches_brute_force(log):
    t = 64
    s = rng(8).random((t, t // 2 + 1)) * 10 + 1
    # box for tile 64: int(64/15) = 4 -> 5 -> clamped to 7 (GRmBackground2D.cu:33-36)
    out = O.rm_background(s, log)
    assert np.allclose(out, _numpy_background(s, 3, bool(log)), atol=1e-10)


def test_background_constant_input_is_zero():
    s = np.full((512, 257), 4.0)
    out = O.rm_background(s)
    assert np.allclose(out, 0.0, atol=1e-12)


# ---------------------------------------------------------------- a12 flatness gate
def test_moments_match_numpy():
    for seed, shape in ((10, (512, 257)), (11, (64, 33)), (12, (7, 5))):
        s = rng(seed).standard_normal(shape) * 3.0 + 1.5
        m = O.moments(s)
        assert m[0] == pytest.approx(s.mean(), rel=1e-13, abs=1e-13)
        assert m[1] == pytest.approx((s * s).mean(), rel=1e-13)
        assert m[2] == pytest.approx(s.var(), rel=1e-10)


def _two_level(mean, var, n=512 * 257):
    """A spectrum of n values mean +- sqrt(var): its mean and variance are
    exactly the given ones up to the FP64 rounding of the square."""
    d = math.sqrt(var)
    s = np.full(n, mean, np.float64)
    s[0::2] += d
    s[1::2] -= d
    return s.reshape(512, 257)


def _gate_f32(s):
    # GCalcMoment2D's two float results and fStd - fMean * fMean in IEEE
    # single, each operation rounded (numpy float32 arithmetic)
    m1, m2 = np.float32(s.mean()), np.float32((s * s).mean())
    return int(np.float32(m2 - np.float32(m1 * m1)) >= np.float32(1.0))


def test_gate_is_the_reference_float_rule():
    """The gate runs the low-pass iff fStd - fMean^2 >= 1 in IEEE single
    (CFindCtfBase.cpp:171-174).  Near mean 4097 the float spacing of the
    squares is 2, so
    the single-precision variance differs from the FP64 one: spectra where
    FP32 and FP64 disagree, in both directions, decide as numpy's float32
    arithmetic does; ordinary spectra decide as their variance says.  Near
    mean 6000 (spacing 4) a variance above 1 can round away entirely."""
    disagree = {0: 0, 1: 0}
    for mean in np.concatenate([np.arange(4090.0, 4100.0, 1.0), np.arange(6000.0, 6010.0, 1.0)]):
        for var in (0.3, 0.7, 0.9, 1.1, 1.3, 1.7, 2.4):
            s = _two_level(mean, var)
            want = _gate_f32(s)
            assert O.gate(s) == want, (mean, var)
            if want != int(var >= 1.0):
                disagree[want] += 1
    assert disagree[0] > 0 and disagree[1] > 0, disagree
    # small means: the float rule and the variance agree
    for var, want in ((0.5, 0), (0.999, 0), (1.001, 1), (50.0, 1)):
        assert O.gate(_two_level(3.0, var)) == want


def test_flat_spectrum_skips_lowpass():
    """A spectrum whose background-removed variance is below 1 is left
    unfiltered (CFindCtfBase.cpp:174 returns before mLowpass)."""
    raw = 1e-3 * (rng(13).random((512, 257)) + 1.0)
    fc = O.FindCtf2D(pix=1.2)
    fc.set_raw_spect(raw)
    assert not fc.lowpassed()
    _, sp, _ = fc.spects()
    assert np.allclose(sp, O.rm_background(raw), rtol=0, atol=0)
    fc.set_raw_spect(raw * 1e4)
    assert fc.lowpassed()
    _, sp, _ = fc.spects()
    assert np.allclose(sp, O.lowpass(O.rm_background(raw * 1e4)), rtol=1e-12, atol=1e-9)


# ---------------------------------------------------------------- a13 low-pass
def test_lowpass_matches_numpy_transfer():
    t = 64
    s = rng(9).random((t, t // 2 + 1))
    full = np.empty((t, t))
    for yd in range(t):
        for xd in range(t):
            xs, ys = xd - t // 2, yd
            if xs < 0:
                xs, ys = -xs, (t - yd) % t
            full[yd, xd] = s[ys, xs]
    k = np.fft.fftfreq(t) * t
    scale = float(np.float32(np.float32(-10.0) / (2.0 * t * t)))
    H = np.exp(scale * (k[:, None] ** 2 + k[None, :] ** 2))
    lp = np.real(np.fft.ifft2(np.fft.fft2(full) * H))
    ref = lp[:, [(t // 2 + c) % t for c in range(t // 2 + 1)]]
    assert np.allclose(O.lowpass(s), ref, atol=1e-12)
    assert np.allclose(O.lowpass(np.full((t, t // 2 + 1), 2.5)), 2.5, atol=1e-12)


# ---------------------------------------------------------------- a14 radial
def test_radial_table_fp32_facts():
    counts, xy = O.radial_table()
    counts64, xy64, dist = O.radial_table(f64=True)
    assert len(xy) == 41016 and np.array_equal(counts, counts64)
    off = np.concatenate([[0], np.cumsum(counts)[:-1]])
    diff = np.nonzero((xy != xy64).any(1))[0]
    where = [(int(np.searchsorted(off, i, side="right") - 1), int(i - off[np.searchsorted(off, i, side="right") - 1]))
             for i in diff]
    assert where == [(1, 2), (102, 148)]          # the two FP32-vs-FP64 samples
    counts_f, xy_f = O.radial_table(use_fma=True)
    assert np.array_equal(xy_f, xy)               # contraction does not change the table
    assert xy[:, 0].min() >= 0 and xy[:, 0].max() <= 256 and xy[:, 1].max() <= 511


def test_radial_avg_constant_and_brute():
    s = rng(10).random((512, 257))
    counts, xy = O.radial_table()
    off = np.concatenate([[0], np.cumsum(counts)])
    ref = np.array([s[xy[off[r]:off[r + 1], 1], xy[off[r]:off[r + 1], 0]].mean() for r in range(257)])
    assert np.allclose(O.radial_avg(s), ref, atol=1e-14)
    assert np.allclose(O.radial_avg(np.full((512, 257), 3.0)), 3.0)


def test_radial_golden_fixture():
    import os
    f = os.path.join(os.path.dirname(__file__), "golden", "radial_table_257x512.npz")
    g = np.load(f)
    counts, xy = O.radial_table()
    assert np.array_equal(g["counts"], counts) and np.array_equal(g["xy"], xy)


# ---------------------------------------------------------------- a15/a18 CTF
def _zero_freq(wl, cs, df, phase, k):
    # CCTFTheory::CalcFrequency (CCTFTheory.cpp:201-228): a s^4 + b s^2 + c = k pi
    a = -0.5 * math.pi * wl ** 3 * cs
    b = math.pi * wl * df
    c = phase - k * math.pi
    det = b * b - 4 * a * c
    s1 = (-b + math.sqrt(det)) / (2 * a)
    s2 = (-b - math.sqrt(det)) / (2 * a)
    return math.sqrt(s1 if s1 > 0 else s2)


def test_ctf1d_zeros():
    p = O.ctf_params(300.0, 2.7, 0.07, 1.2)
    wl, cs, amp = float(p[0]), float(p[1]), float(p[2])
    df = 15000.0 / 1.2
    pi1 = float(np.float32(3.141592654))
    for k in (1, 2, 5, 9):
        s = _zero_freq(wl * pi1 / math.pi, cs, df, amp, k)  # kernel's own pi constant
        # evaluate the 1D CTF on a fine grid by rescaling n so that bin i sits at s
        n = 257
        i = s / (0.5 / (n - 1))
        lo, hi = int(math.floor(i)), int(math.floor(i)) + 1
        c = O.ctf1d(wl, cs, df, amp, n)
        assert c[lo] * c[hi] <= 0                 # sign change brackets the zero


def test_ctf2d_axes_match_1d_and_generator():
    p = O.ctf_params(300.0, 2.7, 0.07, 1.0)
    wl, cs, amp = float(p[0]), float(p[1]), float(p[2])
    c2 = O.ctf2d(wl, cs, 14000.0, 16000.0, 0.0, amp)     # azimuth 0: dfmax along +x
    c1max = O.ctf1d(wl, cs, 16000.0, amp)
    # along +x (row y = 256) the 2D CTF is the 1D CTF at dfmax (pi differs by 1 ulp)
    assert np.allclose(c2[256, :], c1max, atol=2e-4)
    # the generator's independent CTF (k in 1/A, atan2) agrees on the grid
    from gctffind_amd import synth
    g, _ = synth.ctf_rfft(512, 512, 1.0, df1=16000.0, df2=14000.0, az_deg=0.0)
    g = np.roll(g.numpy(), 256, axis=0)
    assert np.allclose(c2[:, :256], g[:, :256], atol=5e-4)


def test_full_spectrum_layout_and_painting():
    """orc_fc_full_spectrum, the -OutMrc diagnostic (CFindCtfBase.cpp:118-133,
    CSpectrumImage.cpp:42-71, GCalcCTF2D.cu:49-78): the right half is the
    half spectrum, the left half its point mirror, and inside the resolution
    ring [25 p, max(2 p / 0.8, 3.5)] A the left half is painted with
    (|CTF| - 1/2) 1.5 sd + mean of the half spectrum at the fitted
    parameters.  The mirror is built with numpy flips, the CTF by the
    synthetic generator's own atan2 formula, the moments by numpy."""
    from gctffind_amd import synth
    img = synth.micrograph_np(1024, 1024, pix=1.2, seed=3, df1=15500.0, df2=14500.0, az_deg=30.0)
    nnx, nny, _, pixn, _ = O.rescale_sizes(1024, 1024, 1.2)
    fc = O.FindCtf2D(pix=pixn)
    fc.gen_half_spectrum(O.rescale(img, 1.2), nnx, nny)
    r = fc.do2d()
    out = fc.full_spectrum()
    h = fc.spects()[1]
    assert out.shape == (512, 512)
    # right half: the half spectrum's columns 0..255 as they are
    assert np.array_equal(out[:, 256:], h[:, :256])
    # left half: the point mirror (y -> -y mod 512, x -> 256 - x)
    mirror = np.roll(np.flip(h, axis=0), 1, axis=0)[:, 256:0:-1]
    y, bx = np.mgrid[0:512, 0:256]
    f = np.sqrt(((bx - 256) * 0.5 / 256.0) ** 2 + ((y - 256.0) / 512.0) ** 2)
    p = np.float32(pixn)
    ring = (f >= p / (25.0 * p)) & (f <= p / max(2.0 * p / 0.8, 3.5))
    assert ring.sum() > 10000 and (~ring).sum() > 10000
    assert np.array_equal(out[:, :256][~ring], mirror[~ring])
    # the painted ring: |CTF| at the point-mirrored frequency (the CTF is even)
    g, _ = synth.ctf_rfft(512, 512, pixn, df1=r["df_max"], df2=r["df_min"], az_deg=r["azimuth"],
                          ext_phase_deg=r["ext_phase"])
    g = np.abs(g.numpy())[(256 - y) % 512, 256 - bx]
    want = (g - 0.5) * 1.5 * h.std() + h.mean()
    assert np.allclose(out[:, :256][ring], want[ring], rtol=0, atol=1e-3 * 1.5 * h.std())


# ---------------------------------------------------------------- a16/a19 CC
def test_cc1d_brute_and_invariance():
    n = 257
    ctf = rng(11).uniform(-1, 1, n)
    spec = rng(12).standard_normal(n)
    lo, hi = 20.48, 175.54
    x = np.arange(n)
    m = (x >= lo) & (x < hi)
    c = (np.abs(ctf) - 0.5) * np.exp(-(x / 256.0) ** 2)
    ref = np.sum(c[m] * spec[m]) / math.sqrt(np.sum(c[m] ** 2) * np.sum(spec[m] ** 2))
    assert O.cc1d(ctf, spec, lo, hi) == pytest.approx(ref, abs=1e-7)
    assert O.cc1d(ctf, 3.0 * spec, lo, hi) == pytest.approx(O.cc1d(ctf, spec, lo, hi), abs=1e-7)


def test_cc2d_brute_and_self():
    ctf = rng(13).uniform(-1, 1, (512, 257))
    spec = rng(14).standard_normal((512, 257))
    flo, fhi = 20.48, 175.54
    y, x = np.mgrid[0:512, 0:257]
    s2 = (x / 512.0) ** 2 + ((y - 256) / 512.0) ** 2
    m = (s2 >= (flo / 512) ** 2) & (s2 <= (fhi / 512) ** 2)
    c = (np.abs(ctf) - 0.5) * np.exp(-16 * s2)
    ref = np.sum(c[m] * spec[m]) / math.sqrt(np.sum(c[m] ** 2) * np.sum(spec[m] ** 2))
    assert O.cc2d(ctf, spec, flo, fhi) == pytest.approx(ref, abs=1e-10)
    assert O.cc2d(ctf, c, flo, fhi) == pytest.approx(1.0, abs=1e-12)   # template itself
    idx = O.ring_pixels(flo, fhi)
    assert len(idx) == int(m.sum()) == 47905


def test_spectral_cc_brute():
    ctf = rng(15).uniform(-1, 1, (64, 33))
    spec = ctf * 2 + rng(16).standard_normal((64, 33)) * 0.5
    shell, cc = O.spectral_cc(ctf, spec, 10)
    y, x = np.mgrid[0:64, 0:33]
    fy = np.abs(y - 32.0)
    r = np.sqrt(x * x + fy * fy)
    for k in (0, 7, 16, 30):
        lo, hi = k - 5, k + 5
        if lo < 0:
            lo, hi = 0, 10
        elif hi > 33:
            lo, hi = 23, 33
        m = (fy < hi) & (x < hi) & (r >= lo) & (r < hi)
        ref = np.corrcoef(ctf[m], spec[m])[0, 1]
        assert cc[k] == pytest.approx(ref, abs=1e-9)


# ---------------------------------------------------------------- a17/a20 search
def _template_spectrum(pix, df1, df2, az_deg):
    p = O.ctf_params(300.0, 2.7, 0.07, pix)
    c = O.ctf2d(float(p[0]), float(p[1]), df2 / pix, df1 / pix, math.radians(az_deg), float(p[2]))
    y, x = np.mgrid[0:512, 0:257]
    s2 = (x / 512.0) ** 2 + ((y - 256) / 512.0) ** 2
    return (np.abs(c) - 0.5) * np.exp(-4 * s2)


def test_search_recovers_noiseless_parameters():
    fc = O.FindCtf2D(pix=1.2)
    fc.trace(True)
    fc.set_half_spect(_template_spectrum(1.2, 21000.0, 20000.0, 35.0))
    fc.set_phase(0.0, 0.0)
    r = fc.do2d()
    assert abs(0.5 * (r["df_min"] + r["df_max"]) - 20500.0) < 30.0
    assert abs((r["df_max"] - r["df_min"]) - 1000.0) < 60.0
    assert abs(r["azimuth"] - 35.0) < 2.0
    tr = fc.get_trace()
    st, cnt = np.unique(tr[:, 0].astype(int), return_counts=True)
    cnt = dict(zip(st, cnt))
    # default 1D grid at 1.2 A: 501 points of 77.76 A over [4320, 43200]
    # (CFindDefocus1D.cpp:78-82); refine: 3000 A / 10 A steps + 1 = 301 or more
    d1 = tr[tr[:, 0] == O.ST_1D]
    assert cnt[O.ST_1D] == 501
    assert d1[0, 1] == pytest.approx(4320.0, abs=0.01) and d1[-1, 1] == pytest.approx(43200.0, abs=0.05)
    assert np.diff(d1[:, 1]).mean() == pytest.approx(77.76, abs=1e-3)
    assert cnt[O.ST_INIT] == 1 and cnt[O.ST_ASTIG] == 2601


def test_search1d_configured_scan():
    """The Do1D scan set by set_search1d: 3000-40000 A at 100 A is 371 points
    (the reference's (int)(range / step) / 2 * 2 + 1 rule), df = lo + i step
    in float; after the refine stage (10 A steps) it lands within one refine
    step of the reference scan's result."""
    fc = O.FindCtf2D(pix=1.2)
    fc.trace(True)
    fc.set_half_spect(_template_spectrum(1.2, 21000.0, 20000.0, 35.0))
    fc.set_phase(0.0, 0.0)
    fc.set_search1d(3000.0, 40000.0, 100.0)
    r = fc.do1d()
    d1 = fc.get_trace()
    d1 = d1[d1[:, 0] == O.ST_1D]
    assert len(d1) == 371
    want = np.float32(3000.0) + np.arange(371, dtype=np.float32) * np.float32(100.0)
    assert np.array_equal(d1[:, 1], want)
    fc.set_search1d()                              # back to the reference scan
    fc.trace(False); fc.trace(True)
    r0 = fc.do1d()
    d1 = fc.get_trace()
    assert (d1[:, 0] == O.ST_1D).sum() == 501
    assert abs(r["df_min"] - r0["df_min"]) <= 10.0


def test_refine_respects_window_caps():
    fc = O.FindCtf2D(pix=1.2)
    fc.trace(True)
    fc.set_half_spect(_template_spectrum(1.2, 12300.0, 11700.0, 50.0))
    r = fc.refine([12000.0, 5000.0], [0.01, 0.04], [50.0, 30.0], [0.0, 0.0])
    tr = fc.get_trace()
    ast = tr[tr[:, 0] == O.ST_ASTIG][:, 2]
    assert ast.max() <= 0.02 + 1e-7                # Refine caps astigmatism (CFindDefocus2D.cpp:141)
    ang = tr[tr[:, 0] == O.ST_ASTIG][:, 3]
    assert ang.min() >= 35.0 - 1e-4 and ang.max() < 65.0
    assert abs(0.5 * (r["df_min"] + r["df_max"]) - 12000.0) < 40.0


@pytest.mark.slow
def test_oracle_recovers_synthetic_micrograph():
    from gctffind_amd import synth
    img = synth.micrograph_np(1024, 1024, pix=1.2, seed=1, df1=15500.0, df2=14500.0, az_deg=30.0)
    r, _ = O.estimate(img, 1.2)
    assert abs(r["df_min"] - 14500.0) < 60 and abs(r["df_max"] - 15500.0) < 60
    assert abs(r["azimuth"] - 30.0) < 2.0
