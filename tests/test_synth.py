"""The synthetic micrograph generator's in-plane defocus gradient (the
tilted-specimen input of the tilt-series workload): without a gradient the
generator is unchanged, and with one, the columns whose defocus offset is
one of the blended offsets carry exactly that offset's CTF-filtered signal."""
import math

import numpy as np
import torch

from gctffind_amd import synth


def _fobj(ny, nx, seed=3):
    g = torch.Generator()
    g.manual_seed(seed)
    return torch.fft.rfft2(torch.randn((ny, nx), generator=g).to(torch.float64))


def test_no_gradient_is_the_single_defocus_signal():
    ny, nx, pix = 96, 128, 1.35
    f = _fobj(ny, nx)
    s, _ = synth.ctf_signal(f, ny, nx, pix, 300.0, 2.7, 0.07, 21000.0, 20400.0, 40.0, 100.0)
    ctf, k2 = synth.ctf_rfft(ny, nx, pix, df1=21000.0, df2=20400.0, az_deg=40.0)
    want = torch.fft.irfft2(f * (ctf * torch.exp(-100.0 * k2 / 4.0)), s=(ny, nx))
    assert torch.equal(s, want)


def test_gradient_end_columns_carry_their_own_defocus():
    ny, nx, pix, tilt = 96, 256, 1.35, 40.0
    ddf = pix * math.tan(math.radians(tilt))             # 1.13 A per pixel: 289 A across, K = 4
    f = _fobj(ny, nx)
    s, _ = synth.ctf_signal(f, ny, nx, pix, 300.0, 2.7, 0.07, 30200.0, 29800.0, 45.0, 100.0, ddf_dx=ddf)
    lo, hi = ddf * (0 - nx / 2), ddf * (nx - 1 - nx / 2)
    for col, off in ((0, lo), (nx - 1, hi)):
        one, _ = synth.ctf_signal(f, ny, nx, pix, 300.0, 2.7, 0.07, 30200.0 + off, 29800.0 + off, 45.0, 100.0)
        assert torch.allclose(s[:, col], one[:, col], rtol=0, atol=1e-12 * float(one.abs().max())), col
    # the defocus really changes across the image: the two end columns differ
    # from a signal at the centre defocus
    mid, _ = synth.ctf_signal(f, ny, nx, pix, 300.0, 2.7, 0.07, 30200.0, 29800.0, 45.0, 100.0)
    assert float((s[:, 0] - mid[:, 0]).abs().max()) > 1e-3 * float(mid.abs().max())


def test_gradient_micrograph_shape_and_seed():
    a = synth.micrograph_np(128, 160, pix=1.35, seed=9, df1=30200.0, df2=29800.0, az_deg=45.0, ddf_dx=0.5)
    b = synth.micrograph_np(128, 160, pix=1.35, seed=9, df1=30200.0, df2=29800.0, az_deg=45.0, ddf_dx=0.5)
    c = synth.micrograph_np(128, 160, pix=1.35, seed=9, df1=30200.0, df2=29800.0, az_deg=45.0)
    assert a.shape == (128, 160) and a.dtype == np.float32 and np.array_equal(a, b)
    assert not np.array_equal(a, c)
