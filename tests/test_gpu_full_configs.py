"""The full-size inputs of the benchmark configurations, checked through
properties that do not depend on the size: a tilt series of 41 micrographs of
3838x3710 at 1.35 A with the tilted specimen's in-plane defocus gradient (the
zero-tilt package equals the single-micrograph full search and the FP64
oracle; three later packages equal the oracle's Refine around the same
centre; every package recovers its generating centre defocus),
and an uneven list of 4096^2 float32 and 5760x4092 float16 micrographs at 0.8
and 1.1 A (the list equals the single-micrograph batches), and an
11520x8184 super-resolution micrograph at 0.5 A (equal to its 2x
Fourier-binned copy and, up to the mirrored azimuth, to its transpose), and
a list of eight of them at 1.1 A that splits into two tile work groups."""
import math

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _setup():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gctffind_amd import gctf
    gctf.load()
    return gctf


def test_tilt_series_3838x3710_full_size():
    gctf = _setup()
    from gctffind_amd import synth
    nx, ny, pix = 3838, 3710, 1.35
    tilts = np.arange(-60.0, 60.0 + 1e-3, 3.0, dtype=np.float32)          # 41 tilts
    n = len(tilts)
    imgs = torch.empty((n, ny, nx), dtype=torch.float32, device="cuda")
    truth = []
    span = []
    for i, t in enumerate(tilts):
        dfm = 30000.0 + 2000.0 * math.sin(math.radians(float(t)))
        # tilted specimen: the defocus changes by pix tan(tilt) A per column
        ddf = pix * math.tan(math.radians(float(t)))
        synth.micrograph(ny, nx, pix=pix, seed=100 + i, df1=dfm + 200.0, df2=dfm - 200.0, az_deg=45.0,
                         device="cuda", out=imgs[i], ddf_dx=ddf)
        truth.append(dfm)
        span.append(abs(ddf) * nx)
    torch.cuda.synchronize()
    ctx = gctf.Context()
    got = ctx.estimate_series(imgs.data_ptr(), n, nx, ny, pix, tilts, tomo=True)
    assert len(got) == n
    i0 = int(np.argmin(np.abs(tilts)))
    one = gctf.Context().estimate_batch(imgs[i0].data_ptr(), 1, nx, ny, pix)[0]
    for k in ("df_min", "df_max", "azimuth", "score"):
        assert got[i0][k] == one[k], (k, got[i0][k], one[k])
    for i in range(n):
        dm = 0.5 * (got[i]["df_min"] + got[i]["df_max"])
        print(f"tilt {float(tilts[i]):6.1f}  defocus span {span[i]:6.0f} A  estimate {dm:8.1f}  truth {truth[i]:8.1f}",
              flush=True)
        # the tile average mixes the defocus values of all columns
        assert abs(dm - truth[i]) < 500.0 + 0.1 * span[i], (i, float(tilts[i]), dm, truth[i])
    ro, _ = O.estimate(imgs[i0].cpu().numpy(), pix)
    assert abs(one["df_min"] - ro["df_min"]) <= 5.0 and abs(one["df_max"] - ro["df_max"]) <= 5.0
    assert abs(one["azimuth"] - ro["azimuth"]) <= 0.5
    assert abs(one["score"] - ro["score"]) <= 1e-4
    # single packages of the chain against the oracle's Refine around the
    # same centre: the chain's reference image before each package (the
    # best score so far, CFindSeriesCtfs.cpp:196-202), taken from the GPU's
    # own results; tomo: astigmatism and azimuth fixed (:188-194)
    order = [i0] + [i for i in range(n) if i != i0]
    ref, center = i0, {}
    for i in order[1:]:
        center[i] = ref
        if got[ref]["score"] < got[i]["score"]:
            ref = i
    for i in (order[1], order[2], n - 1):
        rp = got[center[i]]
        dm = 0.5 * (rp["df_min"] + rp["df_max"])
        ast = (rp["df_max"] - rp["df_min"]) * 0.5 / dm
        fc = O.FindCtf2D(pix=pix)
        fc.gen_half_spectrum(O.rescale(imgs[i].cpu().numpy(), pix), nx, ny)
        r = fc.refine(np.float32([dm, max(dm * 0.5, 5000.0 * pix * pix)]), np.float32([ast, 0.0]),
                      [rp["azimuth"], 0.0], [rp["ext_phase"], 0.0])
        print(f"package {i} (tilt {float(tilts[i])}) around image {center[i]}: gpu {got[i]} oracle {r}", flush=True)
        assert abs(got[i]["df_min"] - r["df_min"]) <= 5.0 and abs(got[i]["df_max"] - r["df_max"]) <= 5.0, i
        assert abs(got[i]["azimuth"] - r["azimuth"]) <= 0.5, i
        assert abs(got[i]["score"] - r["score"]) <= 1e-4, i


def test_uneven_full_size_list():
    gctf = _setup()
    from gctffind_amd import synth
    spec = [  # (nx, ny, pix, dtype, mean defocus, astigmatism, azimuth)
        (4096, 4096, 0.8, gctf.DTYPE_F32, 12000.0, 600.0, 25.0),
        (5760, 4092, 0.8, gctf.DTYPE_F16, 21000.0, 900.0, -40.0),
        (5760, 4092, 1.1, gctf.DTYPE_F16, 33000.0, 400.0, 70.0),
    ]
    bufs, images = [], []
    for i, (nx, ny, pix, dt, dfm, ast, az) in enumerate(spec):
        img = synth.micrograph(ny, nx, pix=pix, seed=200 + i, df1=dfm + ast / 2, df2=dfm - ast / 2, az_deg=az,
                               device="cuda")
        img = img.to(torch.float16 if dt == gctf.DTYPE_F16 else torch.float32).contiguous()
        bufs.append(img)
        images.append(dict(data=img.data_ptr(), dtype=dt, nx=nx, ny=ny, pix=pix))
    torch.cuda.synchronize()
    got = gctf.Context().estimate_list(images)
    for i, im in enumerate(images):
        one = gctf.Context().estimate_batch(im["data"], 1, im["nx"], im["ny"], im["pix"], dtype=im["dtype"])[0]
        for k in ("df_min", "df_max", "azimuth", "score", "ctf_res"):
            assert got[i][k] == one[k], (i, k, got[i][k], one[k])
        dm = 0.5 * (got[i]["df_min"] + got[i]["df_max"])
        assert abs(dm - spec[i][4]) < 500.0, (i, dm, spec[i][4])


def test_super_resolution_11520x8184():
    """An 11520x8184 micrograph at 0.5 A and its 2x Fourier-binned 5760x4092
    copy at 1.0 A rescale to the same 4800x3410 Fourier crop at 1.2 A, so
    their estimates agree; both recover the generating defocus, and in an
    uneven float16 list each equals its single-micrograph batch."""
    gctf = _setup()
    from gctffind_amd import synth
    ny, nx, pix = 8184, 11520, 0.5
    dfm, ast, az = 21000.0, 600.0, 35.0
    big = synth.micrograph(ny, nx, pix=pix, seed=250, df1=dfm + ast / 2, df2=dfm - ast / 2, az_deg=az,
                           device="cuda")
    f = torch.fft.rfft2(big.to(torch.float64))
    h, w = ny // 2, nx // 2
    crop = torch.cat([f[:h // 2, :w // 2 + 1], f[ny - h // 2:, :w // 2 + 1]], dim=0)
    binned = (torch.fft.irfft2(crop, s=(h, w)) / 4.0).to(torch.float32).contiguous()
    del f, crop
    torch.cuda.synchronize()
    ctx = gctf.Context()
    rd = ctx.estimate_batch(big.data_ptr(), 1, nx, ny, pix)[0]
    rb = ctx.estimate_batch(binned.data_ptr(), 1, w, h, 2 * pix)[0]
    for r in (rd, rb):
        assert abs(0.5 * (r["df_min"] + r["df_max"]) - dfm) < 500.0, r
    assert abs(rd["df_min"] - rb["df_min"]) <= 20.0 and abs(rd["df_max"] - rb["df_max"]) <= 20.0, (rd, rb)
    assert abs(rd["azimuth"] - rb["azimuth"]) <= 1.0, (rd, rb)
    # portrait: the transposed micrograph (8184 wide, 11520 high) runs the split
    # column transform; defocus is unchanged and the azimuth mirrors to 90 - az
    rt = ctx.estimate_batch(big.t().contiguous().data_ptr(), 1, ny, nx, pix)[0]
    assert abs(rt["df_min"] - rd["df_min"]) <= 30.0 and abs(rt["df_max"] - rd["df_max"]) <= 30.0, (rt, rd)
    assert abs((90.0 - rd["azimuth"] - rt["azimuth"] + 90.0) % 180.0 - 90.0) <= 2.0, (rt, rd)
    b16, s16 = big.to(torch.float16).contiguous(), binned.to(torch.float16).contiguous()
    images = [dict(data=b16.data_ptr(), dtype=gctf.DTYPE_F16, nx=nx, ny=ny, pix=pix),
              dict(data=s16.data_ptr(), dtype=gctf.DTYPE_F16, nx=w, ny=h, pix=2 * pix)]
    torch.cuda.synchronize()
    got = gctf.Context().estimate_list(images)
    for i, im in enumerate(images):
        one = gctf.Context().estimate_batch(im["data"], 1, im["nx"], im["ny"], im["pix"], dtype=im["dtype"])[0]
        for k in ("df_min", "df_max", "azimuth", "score", "ctf_res"):
            assert got[i][k] == one[k], (i, k, got[i][k], one[k])
        assert abs(0.5 * (got[i]["df_min"] + got[i]["df_max"]) - dfm) < 500.0, (i, got[i])


def test_list_split_into_groups_equals_single_batches():
    """Eight 11520x8184 float16 micrographs at 1.1 A need 1.18 GB of row-pass
    output each, so the tile work list splits into two groups under its
    8 GiB budget; every micrograph still equals its single-micrograph batch."""
    gctf = _setup()
    from gctffind_amd import synth
    ny, nx, pix = 8184, 11520, 1.1
    bufs, images, truth = [], [], []
    for i in range(8):
        dfm = 12000.0 + 3000.0 * i
        img = synth.micrograph(ny, nx, pix=pix, seed=280 + i, df1=dfm + 300.0, df2=dfm - 300.0, az_deg=20.0 * i,
                               device="cuda")
        bufs.append(img.to(torch.float16).contiguous())
        del img
        images.append(dict(data=bufs[-1].data_ptr(), dtype=gctf.DTYPE_F16, nx=nx, ny=ny, pix=pix))
        truth.append(dfm)
    torch.cuda.synchronize()
    got = gctf.Context().estimate_list(images)
    for i, im in enumerate(images):
        one = gctf.Context().estimate_batch(im["data"], 1, nx, ny, pix, dtype=gctf.DTYPE_F16)[0]
        for k in ("df_min", "df_max", "azimuth", "score", "ctf_res"):
            assert got[i][k] == one[k], (i, k, got[i][k], one[k])
        assert abs(0.5 * (got[i]["df_min"] + got[i]["df_max"]) - truth[i]) < 500.0, (i, got[i])
