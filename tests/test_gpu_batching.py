"""Batching entry points: an uneven list of micrographs (shapes, float16 and
float32, pixel sizes) equals the per-micrograph batch results; a tilt series
chained on the device equals the reference's package loop restated with the
oracle; the autotuner's choice leaves results unchanged."""
import math

import numpy as np
import pytest

import oracle as O
from test_gpu_cli import oracle_series

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _setup():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gctffind_amd import gctf
    gctf.load()
    return gctf


def test_uneven_list_equals_single_batches():
    gctf = _setup()
    from gctffind_amd import synth
    spec = [  # (nx, ny, pix, dtype, df1, df2, az)
        (1024, 1024, 1.2, gctf.DTYPE_F32, 15000.0, 14400.0, 20.0),
        (1280, 1024, 0.9, gctf.DTYPE_F16, 21000.0, 20500.0, -35.0),
        (1536, 1536, 1.0, gctf.DTYPE_F32, 9000.0, 8800.0, 60.0),
        (1024, 1200, 1.35, gctf.DTYPE_F16, 26000.0, 25000.0, -80.0),
        # same shapes again: rescale classes with several micrographs
        (1536, 1536, 1.0, gctf.DTYPE_F32, 12500.0, 12000.0, 5.0),
        (1280, 1024, 0.9, gctf.DTYPE_F16, 17000.0, 16200.0, 75.0),
    ]
    bufs, images = [], []
    for i, (nx, ny, pix, dt, a, b, az) in enumerate(spec):
        img = synth.micrograph(ny, nx, pix=pix, seed=400 + i, df1=a, df2=b, az_deg=az, device="cuda")
        img = img.to(torch.float16 if dt == gctf.DTYPE_F16 else torch.float32).contiguous()
        bufs.append(img)
        images.append(dict(data=img.data_ptr(), dtype=dt, nx=nx, ny=ny, pix=pix))
    torch.cuda.synchronize()
    ctx = gctf.Context()
    got = ctx.estimate_list(images)
    got2 = ctx.estimate_list(images)   # graph replay
    cg = gctf.Context()
    cg.tune(list_queue=-1)             # tile row pass with one workgroup per item
    got3 = cg.estimate_list(images)
    for i, im in enumerate(images):
        one = gctf.Context().estimate_batch(im["data"], 1, im["nx"], im["ny"], im["pix"], dtype=im["dtype"])[0]
        for k in ("df_min", "df_max", "azimuth", "score", "ctf_res"):
            assert got[i][k] == one[k], (i, k, got[i][k], one[k])
            assert got2[i][k] == one[k], (i, k)
            assert got3[i][k] == one[k], (i, k)


def test_search_blocking_keeps_results():
    """2, 4, 8 or 16 candidates per 2D-evaluator workgroup: identical results
    (each candidate's sums are formed on their own)."""
    gctf = _setup()
    from gctffind_amd import synth
    n = 3
    imgs = torch.empty((n, 1024, 1024), dtype=torch.float32, device="cuda")
    for i in range(n):
        synth.micrograph(1024, 1024, pix=1.2, seed=450 + i, df1=10000.0 + 6000 * i + 300, df2=10000.0 + 6000 * i - 300,
                         az_deg=-50.0 + 40 * i, device="cuda", out=imgs[i])
    torch.cuda.synchronize()
    res = {}
    for cb in (2, 4, 8, 16):
        ctx = gctf.Context()
        ctx.tune(search_block=cb)
        res[cb] = ctx.estimate_batch(imgs.data_ptr(), n, 1024, 1024, 1.2)
    for cb in (2, 4, 16):
        for i in range(n):
            for k in ("df_min", "df_max", "azimuth", "score", "ctf_res"):
                assert res[cb][i][k] == res[8][i][k], (cb, i, k)


@pytest.mark.parametrize("tomo", [True, False])
def test_series_matches_oracle_package_loop(tomo):
    """Seven 1024^2 images at 1.35 A (no rescale), tilts -9..9; tomo: the
    reference's tilt-series refine (astigmatism and azimuth fixed)."""
    gctf = _setup()
    from gctffind_amd import synth
    pix = 1.35
    tilts = np.array([-9.0, -6.0, -3.0, 0.0, 3.0, 6.0, 9.0], np.float32)
    imgs = []
    for k, t in enumerate(tilts):
        df = 30000.0 + 2000.0 * math.sin(math.radians(t))
        imgs.append(synth.micrograph_np(1024, 1024, pix=pix, seed=500 + k, df1=df + 200.0, df2=df - 200.0,
                                        az_deg=45.0))
    d = torch.from_numpy(np.stack(imgs)).cuda()
    ctx = gctf.Context()
    got = ctx.estimate_series(d.data_ptr(), len(imgs), 1024, 1024, pix, tilts, tomo=tomo)
    if tomo:
        ref = oracle_series(imgs, pix, tilts)
    else:
        ref = _oracle_series_single(imgs, pix)
    for i in range(len(imgs)):
        assert abs(got[i]["df_min"] - ref[i]["df_min"]) <= 5.0, (i, got[i], ref[i])
        assert abs(got[i]["df_max"] - ref[i]["df_max"]) <= 5.0, (i, got[i], ref[i])
        assert abs(got[i]["azimuth"] - ref[i]["azimuth"]) <= 0.5, (i, got[i], ref[i])
        assert abs(got[i]["score"] - ref[i]["score"]) <= 1e-4, (i, got[i], ref[i])


def test_series_reference_changes_match_oracle_package_loop():
    """A tilt series whose later files score higher than the zero-tilt
    package, so the reference moves (to files 4, 5 and 6 in processing
    order): the batched refine rounds must settle each package against the
    reference the sequential package loop would have used.  In tomo mode
    astigmatism and azimuth stay at the reference's values, so a stale
    reference would show in dfMin / dfMax."""
    gctf = _setup()
    from gctffind_amd import synth
    pix = 1.35
    tilts = np.array([-9.0, -6.0, -3.0, 0.0, 3.0, 6.0, 9.0], np.float32)
    snr = [0.02, 0.03, 0.04, 0.08, 0.2, 0.4, 0.8]
    imgs = []
    for k, t in enumerate(tilts):
        df = 20000.0 + 1500.0 * math.sin(math.radians(t))
        ast = 150.0 + 60.0 * k
        imgs.append(synth.micrograph_np(1024, 1024, pix=pix, seed=700 + k, df1=df + ast, df2=df - ast,
                                        az_deg=20.0 + 7.0 * k, snr=snr[k]))
    d = torch.from_numpy(np.stack(imgs)).cuda()
    got = gctf.Context().estimate_series(d.data_ptr(), len(imgs), 1024, 1024, pix, tilts, tomo=True)
    ref = oracle_series(imgs, pix, tilts)
    for i in range(len(imgs)):
        assert abs(got[i]["df_min"] - ref[i]["df_min"]) <= 5.0, (i, got[i], ref[i])
        assert abs(got[i]["df_max"] - ref[i]["df_max"]) <= 5.0, (i, got[i], ref[i])
        assert abs(got[i]["azimuth"] - ref[i]["azimuth"]) <= 0.5, (i, got[i], ref[i])
        assert abs(got[i]["score"] - ref[i]["score"]) <= 1e-4, (i, got[i], ref[i])
    # the scores rise after the zero-tilt package, so the reference did move
    assert got[4]["score"] > got[3]["score"] and got[6]["score"] > got[5]["score"] > got[4]["score"]


def _oracle_series_single(imgs, pix):
    """CFindSeriesCtfs for single-image packages in one series (not tomo):
    Refine also searches astigmatism (+-0.02) and azimuth (+-15 deg)."""
    n = len(imgs)
    res, ref = [None] * n, None
    for i in range(n):
        nnx, nny, pitch, pixn, _ = O.rescale_sizes(imgs[i].shape[1], imgs[i].shape[0], pix)
        fc = O.FindCtf2D(pix=pixn)
        fc.gen_half_spectrum(O.rescale(imgs[i], pix), nnx, nny)

        def full():
            fc.set_phase(0.0, 0.0)
            return fc.do2d()
        if i == 0:
            r, ref = full(), 0
        else:
            rp = res[ref]
            dm = 0.5 * (rp["df_min"] + rp["df_max"])
            mean = (rp["df_max"] + rp["df_min"]) * 0.5
            ast = 0.0 if mean <= 0 else (rp["df_max"] - rp["df_min"]) * 0.5 / mean
            r = fc.refine(np.float32([dm, max(dm * 0.5, 5000.0 * pix * pix)]), np.float32([ast, 0.04]),
                          [rp["azimuth"], 30.0], [rp["ext_phase"], 0.0])
            if res[ref]["score"] < r["score"]:
                ref = i
            if r["score"] <= 0:
                r, ref = full(), i
        res[i] = r
    return res


def test_autotune_keeps_results():
    gctf = _setup()
    from gctffind_amd import synth
    n, pix = 6, 1.0
    imgs = torch.empty((n, 1536, 1536), dtype=torch.float32, device="cuda")
    for i in range(n):
        synth.micrograph(1536, 1536, pix=pix, seed=600 + i, df1=12000.0 + 2500 * i, df2=11600.0 + 2500 * i,
                         az_deg=-30.0 + 10 * i, device="cuda", out=imgs[i])
    ctx = gctf.Context()
    before = ctx.estimate_batch(imgs.data_ptr(), n, 1536, 1536, pix)
    best = ctx.autotune(imgs.data_ptr(), n, 1536, 1536, pix)
    assert best["chunk"] in (1, 2, 4) and best["tile_splits"] in (4, 8, 16) and best["ms"] > 0
    assert best["search_block"] in (4, 8, 16)
    after = ctx.estimate_batch(imgs.data_ptr(), n, 1536, 1536, pix)
    # the chunk size does not touch the arithmetic; the tile-split count
    # regroups the FP32 tile sums, so results agree to the parity tolerance
    for a, b in zip(before, after):
        assert abs(a["df_min"] - b["df_min"]) <= 5.0 and abs(a["df_max"] - b["df_max"]) <= 5.0
        assert abs(a["azimuth"] - b["azimuth"]) <= 0.5 and abs(a["score"] - b["score"]) <= 1e-4
        if best["tile_splits"] == 8:
            for k in ("df_min", "df_max", "azimuth", "score", "ctf_res"):
                assert a[k] == b[k]
    # a list of micrographs of the tuned shape takes the tuned tile splits:
    # it equals the tuned batch bit for bit
    lst = ctx.estimate_list([dict(data=imgs[i].data_ptr(), dtype=gctf.DTYPE_F32, nx=1536, ny=1536, pix=pix)
                             for i in range(n)])
    for a, b in zip(lst, after):
        for k in ("df_min", "df_max", "azimuth", "score", "ctf_res"):
            assert a[k] == b[k], k
