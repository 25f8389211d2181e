"""The full-size inputs of the benchmark configurations, checked through
properties that do not depend on the size: a tilt series of 41 micrographs of
3838x3710 at 1.35 A (the zero-tilt package equals the single-micrograph full
search and the FP64 oracle; every package recovers its generating defocus),
and an uneven list of 4096^2 float32 and 5760x4092 float16 micrographs at 0.8
and 1.1 A (the list equals the single-micrograph batches)."""
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
    for i, t in enumerate(tilts):
        dfm = 30000.0 + 2000.0 * math.sin(math.radians(float(t)))
        synth.micrograph(ny, nx, pix=pix, seed=100 + i, df1=dfm + 200.0, df2=dfm - 200.0, az_deg=45.0,
                         device="cuda", out=imgs[i])
        truth.append(dfm)
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
        assert abs(dm - truth[i]) < 500.0, (i, float(tilts[i]), dm, truth[i])
    ro, _ = O.estimate(imgs[i0].cpu().numpy(), pix)
    assert abs(one["df_min"] - ro["df_min"]) <= 5.0 and abs(one["df_max"] - ro["df_max"]) <= 5.0
    assert abs(one["azimuth"] - ro["azimuth"]) <= 0.5
    assert abs(one["score"] - ro["score"]) <= 1e-4


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
