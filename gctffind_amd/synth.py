"""Seeded synthetic micrographs (test and benchmark input, not the product path).

Image model (one micrograph):
  signal = irfft2( rfft2(object) * CTF(k) * exp(-B k^2 / 4) )
  background = irfft2( rfft2(white2) * sqrt(1 / (1 + (k / 0.02)^2)) ), scaled so
               var(background) = var(signal)
  noise = white Gaussian with var(noise) = var(signal) / snr
  image = mean + signal + background + noise
object, white2 and noise are independent standard normal fields from a torch
generator seeded with `seed`.  k is in 1/A.

A tilted specimen (ddf_dx != 0) has a linear in-plane defocus gradient: the
defocus at column x is df + ddf_dx (x - nx/2) A (ddf_dx = pix tan(tilt) for
a tilt axis along y).  The signal is then the column-wise linear blend of
signals filtered at K evenly spaced defocus offsets no more than 100 A
apart (each column mixes the two offsets around its own), so every tile of
the micrograph sees the defocus of its own columns.  The CTF follows the
reference's convention (FindCTF/GCalcCTF2D.cu:23-37): the astigmatism azimuth
is the direction of the larger defocus, measured from +x (columns) towards +y
(rows), chi = pi lambda k^2 (df - lambda^2 k^2 Cs / 2) + amplitude phase +
extra phase, CTF = -sin(chi).
"""
import math

import numpy as np
import torch


def wavelength(kv):
    """Electron wavelength in A (FindCTF/CCTFTheory.cpp:283-287)."""
    return 12.26 / math.sqrt(kv * 1000.0 + 0.9784 * kv * kv)


def ctf_rfft(ny, nx, pix, kv=300.0, cs_mm=2.7, ac=0.07, df1=15500.0, df2=14500.0, az_deg=30.0,
             ext_phase_deg=0.0, device="cpu", dtype=torch.float64):
    """CTF sampled on the rfft2 grid of an ny x nx image (float64)."""
    ky = torch.fft.fftfreq(ny, d=pix, device=device, dtype=dtype)[:, None]
    kx = torch.fft.rfftfreq(nx, d=pix, device=device, dtype=dtype)[None, :]
    k2 = kx * kx + ky * ky
    theta = torch.atan2(ky, kx)
    dfmax, dfmin = max(df1, df2), min(df1, df2)
    df = 0.5 * (dfmax + dfmin) + 0.5 * (dfmax - dfmin) * torch.cos(2.0 * (theta - math.radians(az_deg)))
    lam = wavelength(kv)
    cs = cs_mm * 1e7
    amp = math.atan(ac / math.sqrt(1.0 - ac * ac))
    chi = math.pi * lam * k2 * (df - 0.5 * lam * lam * k2 * cs) + amp + math.radians(ext_phase_deg)
    return -torch.sin(chi), k2


GRADIENT_STEP_A = 100.0


def ctf_signal(fobj, ny, nx, pix, kv, cs_mm, ac, df1, df2, az_deg, bfac, ext_phase_deg=0.0, ddf_dx=0.0,
               device="cpu"):
    """The object's spectrum fobj (rfft2 grid) through CTF x envelope, back
    in real space (float64), and k^2 on the rfft2 grid.  With ddf_dx the
    defocus offsets lo + j step (j < K, step <= GRADIENT_STEP_A) are blended
    per column with hat weights, so column x of offset ddf_dx (x - nx/2)
    equal to an offset takes that offset's signal alone."""
    if ddf_dx == 0.0:
        ctf, k2 = ctf_rfft(ny, nx, pix, kv, cs_mm, ac, df1, df2, az_deg, ext_phase_deg, device)
        env = torch.exp(-bfac * k2 / 4.0)
        return torch.fft.irfft2(fobj * (ctf * env), s=(ny, nx)), k2
    off = ddf_dx * (torch.arange(nx, device=device, dtype=torch.float64) - nx / 2)
    lo, hi = float(off.min()), float(off.max())
    k = max(2, int(math.ceil((hi - lo) / GRADIENT_STEP_A)) + 1)
    step = (hi - lo) / (k - 1)
    signal = torch.zeros((ny, nx), device=device, dtype=torch.float64)
    for j in range(k):
        oj = lo + j * step
        ctf, k2 = ctf_rfft(ny, nx, pix, kv, cs_mm, ac, df1 + oj, df2 + oj, az_deg, ext_phase_deg, device)
        env = torch.exp(-bfac * k2 / 4.0)
        w = torch.clamp(1.0 - torch.abs(off - oj) / step, min=0.0)   # hat weights, sum 1 per column
        signal += torch.fft.irfft2(fobj * (ctf * env), s=(ny, nx)) * w[None, :]
        del ctf
    return signal, k2


def micrograph(ny=4096, nx=4096, pix=1.0, seed=0, kv=300.0, cs_mm=2.7, ac=0.07, df1=15500.0,
               df2=14500.0, az_deg=30.0, bfac=100.0, snr=0.1, mean=0.0, ext_phase_deg=0.0,
               device="cpu", out=None, ddf_dx=0.0):
    """One synthetic micrograph as float32 (torch tensor on `device`).
    ddf_dx: in-plane defocus gradient along x, A per pixel (0: none)."""
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    obj = torch.randn((ny, nx), generator=g, device=device, dtype=torch.float32)
    fobj = torch.fft.rfft2(obj.to(torch.float64))
    del obj
    signal, k2 = ctf_signal(fobj, ny, nx, pix, kv, cs_mm, ac, df1, df2, az_deg, bfac, ext_phase_deg, ddf_dx,
                            device)
    del fobj
    w2 = torch.randn((ny, nx), generator=g, device=device, dtype=torch.float32).to(torch.float64)
    bg = torch.fft.irfft2(torch.fft.rfft2(w2) * torch.sqrt(1.0 / (1.0 + k2 / (0.02 ** 2))), s=(ny, nx))
    del w2
    vs = signal.var()
    bg = bg * torch.sqrt(vs / bg.var())
    noise = torch.randn((ny, nx), generator=g, device=device, dtype=torch.float32).to(torch.float64)
    img = mean + signal + bg + noise * torch.sqrt(vs / snr)
    img = img.to(torch.float32)
    if out is not None:
        out.copy_(img)
        return out
    return img


def micrograph_np(*args, **kw):
    """numpy float32 copy of micrograph(...) generated on the CPU."""
    kw["device"] = "cpu"
    return micrograph(*args, **kw).numpy()


def random_params(rng, pix_lo=0.5, pix_hi=1.1, df_lo=5000.0, df_hi=50000.0, ast_max=1500.0):
    """Draw (pix, df1, df2, azimuth) with log-uniform defocus and uniform
    astigmatism and azimuth."""
    pix = float(rng.uniform(pix_lo, pix_hi))
    dfm = float(math.exp(rng.uniform(math.log(df_lo), math.log(df_hi))))
    ast = float(rng.uniform(0.0, ast_max))
    az = float(rng.uniform(0.0, 180.0))
    return pix, dfm + 0.5 * ast, dfm - 0.5 * ast, az


__all__ = ["wavelength", "ctf_rfft", "ctf_signal", "micrograph", "micrograph_np", "random_params", "np"]
