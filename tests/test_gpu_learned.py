"""The trained proposal network against the method without it (the plain
reference search), on synthetic 4096^2 micrographs at 1.0 A (rescaled to
3412^2 at 1.2 A, as the benchmark's workload).

- Outside the reference's 1D search range (3000 p^2 .. 30000 p^2 = 4320 ..
  43200 A at p = 1.2 A, and the 0.9 .. 1.1 window of the 2D search around
  it), the plain search cannot reach the defocus; the learned path, whose
  window follows the proposal, recovers it.
- Inside the range the two agree to 20 A below 24000 A.
- The learned path is not slower than the plain one on a batch.
"""
import math
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _setup():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gctffind_amd import gctf, proposal
    gctf.load()
    ctx = gctf.Context()
    params, _ = proposal.load_params()
    ctx.set_proposal_params(params)
    return gctf, ctx


def _batch(dfs, seed0, ast=300.0, bfac=100.0, snr=0.1):
    from gctffind_amd import synth
    imgs = torch.empty((len(dfs), 4096, 4096), dtype=torch.float32, device="cuda")
    rng = np.random.default_rng(seed0)
    for i, d in enumerate(dfs):
        az = float(rng.uniform(-90, 90))
        synth.micrograph(4096, 4096, pix=1.0, seed=seed0 + i, df1=d + ast / 2, df2=d - ast / 2, az_deg=az,
                         bfac=bfac, snr=snr, device="cuda", out=imgs[i])
    torch.cuda.synchronize()
    return imgs


def _mean(r):
    return 0.5 * (r["df_min"] + r["df_max"])


def test_learned_recovers_defocus_outside_the_reference_range():
    gctf, ctx = _setup()
    dfs = [3100.0, 3300.0, 3500.0, 3700.0, 49000.0, 49800.0]
    imgs = _batch(dfs, 7000)
    n = len(dfs)
    learned, props = ctx.estimate_batch_learned(imgs.data_ptr(), n, 4096, 4096, 1.0, proposals=True)
    plain = ctx.estimate_batch(imgs.data_ptr(), n, 4096, 4096, 1.0)
    el = [abs(_mean(r) - d) for r, d in zip(learned, dfs)]
    ep = [abs(_mean(r) - d) for r, d in zip(plain, dfs)]
    assert max(el) <= 100.0, (el, props)
    assert sum(e > 300.0 for e in ep) >= n - 1, ep


def test_learned_agrees_with_plain_inside_the_range():
    gctf, ctx = _setup()
    rng = np.random.default_rng(11)
    dfs = [float(math.exp(rng.uniform(math.log(5000.0), math.log(20000.0)))) for _ in range(16)]
    imgs = _batch(dfs, 7100)
    n = len(dfs)
    learned, props = ctx.estimate_batch_learned(imgs.data_ptr(), n, 4096, 4096, 1.0, proposals=True)
    plain = ctx.estimate_batch(imgs.data_ptr(), n, 4096, 4096, 1.0)
    d = [abs(_mean(a) - _mean(b)) for a, b in zip(learned, plain)]
    assert sum(x <= 20.0 for x in d) >= n - 1, (d, props)
    # confident proposals, inside the +-300 A window of the truth
    assert all(p["confidence"] >= 0.8 for p in props)
    assert all(abs(p["df_mean"] - t) <= 300.0 for p, t in zip(props, dfs))


def test_learned_not_slower_than_plain():
    gctf, ctx = _setup()
    rng = np.random.default_rng(12)
    dfs = [float(math.exp(rng.uniform(math.log(6000.0), math.log(30000.0)))) for _ in range(32)]
    imgs = _batch(dfs, 7200)
    n = len(dfs)

    def clock(f):
        f()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(3):
            f()
        torch.cuda.synchronize()
        return time.perf_counter() - t
    tp = clock(lambda: ctx.estimate_batch(imgs.data_ptr(), n, 4096, 4096, 1.0))
    tl = clock(lambda: ctx.estimate_batch_learned(imgs.data_ptr(), n, 4096, 4096, 1.0))
    assert tl <= 1.02 * tp, (tl, tp)
