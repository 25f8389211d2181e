"""BF16 MFMA grid search against the FP64 oracle evaluating the same grid.

The GEMM's BF16 correlations are compared loosely (BF16 operands), the best
candidate exactly, and the FP32 re-score of the top 16 within 1e-5 of the
oracle's FP64 CC of the same candidates.
"""
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


def _spectrum(gctf, img, pix):
    """Background-removed, low-passed spectrum of one micrograph from the
    single-image path (the batch path computes the same spectrum)."""
    ctx = gctf.Context()
    nnx, nny, pitch, pixn = ctx.rescale_setup(img.shape[1], img.shape[0], pix)
    d = torch.from_numpy(img).cuda()
    ctx.rescale(d.data_ptr(), on_device=True)
    ctx.setup1(300.0, 2.7, 0.07, pixn)
    ctx.setup2(nnx, nny)
    ctx.gen_half_spectrum(ctx.scaled_image())
    return ctx.get_half_spect(False), pixn


def _cand(grid, m):
    f = np.float32
    n_df, n_ast, n_az = grid["n_df"], grid["n_ast"], grid["n_az"]
    iaz, rest = m % n_az, m // n_az
    iast, idf = rest % n_ast, rest // n_ast
    dstep = (f(grid["df_hi"]) - f(grid["df_lo"])) / f(n_df - 1)
    astep = (f(grid["ast_hi"]) - f(grid["ast_lo"])) / f(n_ast - 1)
    zstep = (f(grid["az_hi"]) - f(grid["az_lo"])) / f(n_az)
    return (f(grid["df_lo"]) + f(idf) * dstep, f(grid["ast_lo"]) + f(iast) * astep,
            f(grid["az_lo"]) + f(iaz) * zstep)


def test_grid_small_matches_oracle():
    gctf = _setup()
    from gctffind_amd import synth
    pix = 1.2
    truths = [(18000.0, 17200.0, 25.0), (12500.0, 12000.0, -50.0), (23000.0, 22000.0, 80.0)]
    imgs = np.stack([synth.micrograph_np(1024, 1024, pix=pix, seed=40 + i, df1=a, df2=b, az_deg=c)
                     for i, (a, b, c) in enumerate(truths)])
    grid = dict(df_lo=11000.0, df_hi=24000.0, n_df=53, ast_lo=0.0, ast_hi=0.06, n_ast=7, az_lo=-90.0,
                az_hi=90.0, n_az=12, res_lo=30.0, res_hi=4.0)
    ctx = gctf.Context()
    d = torch.from_numpy(imgs).cuda()
    res, cc, top, rs = ctx.grid_search_batch(d.data_ptr(), 3, 1024, 1024, pix, grid, debug=True)
    for i in range(3):
        sp, pixn = _spectrum(gctf, imgs[i], pix)
        ref = O.grid_cc(sp.astype(np.float64), grid, pixn)
        # BF16 GEMM correlations track the FP64 ones
        assert np.abs(cc[i] - ref).max() < 0.02
        assert np.corrcoef(cc[i], ref)[0, 1] > 0.999
        best = int(np.argmax(ref))
        assert best in set(top[i].tolist())
        # FP32 re-score of the 16 survivors equals FP64 within 1e-5
        for j, m in enumerate(top[i]):
            assert abs(rs[i, j] - ref[m]) <= 1e-5, (i, j, m, rs[i, j], ref[m])
        dfm, ast, az = _cand(grid, best)
        assert res[i]["df_min"] == pytest.approx(float(dfm * (np.float32(1) - ast)), rel=0, abs=1e-3)
        assert res[i]["df_max"] == pytest.approx(float(dfm * (np.float32(1) + ast)), rel=0, abs=1e-3)
        assert res[i]["azimuth"] == pytest.approx(float(az), abs=1e-4)
        assert res[i]["score"] == pytest.approx(ref[best], abs=1e-5)


def test_grid_full_size_recovers_truth():
    """The 256 x 32 x 36 grid on 4096^2 micrographs at 1.0 A (rescaled to
    1.2 A): best candidates land on the grid point nearest the truth and
    re-score to the oracle's FP64 CC."""
    gctf = _setup()
    from gctffind_amd import synth
    pix = 1.0
    grid = dict(gctf.GRID_256x32x36)
    truths = [(15500.0, 14500.0, 30.0), (26000.0, 25400.0, -20.0)]
    imgs = torch.empty((2, 4096, 4096), dtype=torch.float32, device="cuda")
    for i, (a, b, c) in enumerate(truths):
        synth.micrograph(4096, 4096, pix=pix, seed=60 + i, df1=a, df2=b, az_deg=c, device="cuda", out=imgs[i])
    ctx = gctf.Context()
    res, cc, top, rs = ctx.grid_search_batch(imgs.data_ptr(), 2, 4096, 4096, pix, grid, debug=True)
    step = (grid["df_hi"] - grid["df_lo"]) / (grid["n_df"] - 1)
    for i, (a, b, c) in enumerate(truths):
        dm = 0.5 * (res[i]["df_min"] + res[i]["df_max"])
        assert abs(dm - 0.5 * (a + b)) <= step
        assert rs[i].max() == pytest.approx(res[i]["score"], abs=0)
        sp, pixn = _spectrum(gctf, imgs[i].cpu().numpy(), pix)
        # oracle FP64 CC of the 16 survivors (sub-grids of one candidate each)
        for j, m in enumerate(top[i][:4]):
            dfm, ast, az = _cand(grid, int(m))
            one = dict(df_lo=float(dfm), df_hi=float(dfm), n_df=1, ast_lo=float(ast), ast_hi=float(ast),
                       n_ast=1, az_lo=float(az), az_hi=float(az) + 1.0, n_az=1, res_lo=30.0, res_hi=4.0)
            ref = O.grid_cc(sp.astype(np.float64), one, pixn)[0]
            assert abs(rs[i, j] - ref) <= 1e-5


@pytest.mark.parametrize("n,cols", [(256, (0, 31, 32, 99, 128, 255)), (100, (0, 31, 32, 63, 64, 99)),
                                    (40, (0, 31, 32, 39))])
def test_grid_batch_sizes_match_oracle(n, cols):
    """The GEMM's accumulator width follows the batch: n micrographs make
    ceil(n / 32) column tiles per wave (k_grid_gemm<NT>).  n = 256 (the
    256-micrograph workload) runs NT = 8, n = 100 NT = 4, n = 40 NT = 2.
    Columns in different 32-column tiles are checked against the FP64
    oracle on the same candidates, with the bounds of the small test."""
    gctf = _setup()
    from gctffind_amd import synth
    pix = 1.2
    grid = dict(df_lo=8000.0, df_hi=30000.0, n_df=45, ast_lo=0.0, ast_hi=0.06, n_ast=5, az_lo=-90.0,
                az_hi=90.0, n_az=8, res_lo=30.0, res_hi=4.0)
    rng = np.random.default_rng(n)
    imgs = torch.empty((n, 1024, 1024), dtype=torch.float32, device="cuda")
    for i in range(n):
        dfm = float(rng.uniform(9000.0, 28000.0))
        ast = float(rng.uniform(0.0, 900.0))
        synth.micrograph(1024, 1024, pix=pix, seed=3000 + i, df1=dfm + ast / 2, df2=dfm - ast / 2,
                         az_deg=float(rng.uniform(-90.0, 90.0)), device="cuda", out=imgs[i])
    torch.cuda.synchronize()
    ctx = gctf.Context()
    res, cc, top, rs = ctx.grid_search_batch(imgs.data_ptr(), n, 1024, 1024, pix, grid, debug=True)
    print(f"grid batch n={n} columns {cols}", flush=True)
    for i in cols:
        sp, pixn = _spectrum(gctf, imgs[i].cpu().numpy(), pix)
        ref = O.grid_cc(sp.astype(np.float64), grid, pixn)
        assert np.abs(cc[i] - ref).max() < 0.02, i
        assert np.corrcoef(cc[i], ref)[0, 1] > 0.999, i
        best = int(np.argmax(ref))
        assert best in set(top[i].tolist()), (i, best, top[i])
        for j, m in enumerate(top[i]):
            assert abs(rs[i, j] - ref[m]) <= 1e-5, (i, j, m, rs[i, j], ref[m])
        assert res[i]["score"] == pytest.approx(ref[best], abs=1e-5), i
