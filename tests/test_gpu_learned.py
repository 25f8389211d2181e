"""The trained proposal network against the method without it (the plain
reference search), on synthetic 4096^2 micrographs at 1.0 A (rescaled to
3412^2 at 1.2 A, as the benchmark's workload).

- Outside the reference's 1D search range (3000 p^2 .. 30000 p^2 = 4320 ..
  43200 A at p = 1.2 A, and the 0.9 .. 1.1 window of the 2D search around
  it), the plain search cannot reach the defocus; the narrowed search around
  the proposal recovers it (what the network gains).  The confidence, trained
  on agreement with the plain search, is low there, so at the calibrated
  threshold those micrographs take the plain search's result, as the
  reference would.
- Inside the range the two agree to 20 A below 20000 A; micrographs that fall
  back equal the plain result bit for bit.
- In a batch where only some micrographs fall back, each one equals its own
  single path: the plain search's result, or the narrowed search's.
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


KEYS = ("df_min", "df_max", "azimuth", "score", "ctf_res")


def test_learned_recovers_defocus_outside_the_reference_range():
    gctf, ctx = _setup()
    dfs = [3100.0, 3300.0, 3500.0, 3700.0, 49000.0, 49800.0]
    imgs = _batch(dfs, 7000)
    n = len(dfs)
    # the narrowed search around the proposal, with no confidence fallback
    learned, props = ctx.estimate_batch_learned(imgs.data_ptr(), n, 4096, 4096, 1.0, conf_threshold=0.0,
                                                proposals=True)
    plain = ctx.estimate_batch(imgs.data_ptr(), n, 4096, 4096, 1.0)
    el = [abs(_mean(r) - d) for r, d in zip(learned, dfs)]
    ep = [abs(_mean(r) - d) for r, d in zip(plain, dfs)]
    assert max(el) <= 100.0, (el, props)
    assert sum(e > 300.0 for e in ep) >= n - 1, ep
    # at the calibrated threshold the confidence sends them to the plain search
    cal, cprops = ctx.estimate_batch_learned(imgs.data_ptr(), n, 4096, 4096, 1.0, proposals=True)
    print("confidence", [round(p["confidence"], 3) for p in cprops], flush=True)
    assert sum(p["fallback"] for p in cprops) >= n - 1, cprops
    for a, b, p in zip(cal, plain, cprops):
        if p["fallback"]:
            for k in KEYS:
                assert a[k] == b[k], (k, a, b)


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
    # proposals inside the +-300 A window of the truth
    assert all(abs(p["df_mean"] - t) <= 300.0 for p, t in zip(props, dfs))
    for a, b, p in zip(learned, plain, props):
        if p["fallback"]:
            for k in KEYS:
                assert a[k] == b[k], (k, a, b)


def test_mixed_fallback_batch_equals_single_paths():
    """A threshold between the batch's confidences: the micrographs below it
    take the plain search (equal to gctf_estimate_batch bit for bit), the
    others keep the narrowed search's result (equal to the learned path with
    no confidence fallback)."""
    gctf, ctx = _setup()
    dfs = [6000.0, 9000.0, 14000.0, 19000.0, 24000.0, 30000.0, 36000.0, 42000.0]
    imgs = _batch(dfs, 7300)
    n = len(dfs)
    plain = ctx.estimate_batch(imgs.data_ptr(), n, 4096, 4096, 1.0)
    free, fprops = ctx.estimate_batch_learned(imgs.data_ptr(), n, 4096, 4096, 1.0, conf_threshold=0.0,
                                              proposals=True)
    conf = sorted(p["confidence"] for p in fprops if not p["fallback"])
    print("confidence", [round(p["confidence"], 4) for p in fprops], flush=True)
    # a threshold strictly between two distinct confidences near the middle
    cands = [0.5 * (a + b) for a, b in zip(conf, conf[1:]) if b > a]
    assert cands, conf
    thr = min(cands, key=lambda t: abs(sum(c < t for c in conf) - len(conf) / 2))
    mixed, mprops = ctx.estimate_batch_learned(imgs.data_ptr(), n, 4096, 4096, 1.0, conf_threshold=thr,
                                               proposals=True)
    fb = [p["fallback"] for p in mprops]
    assert 0 < sum(fb) < n, (thr, fb)
    for i in range(n):
        want = plain[i] if fb[i] else free[i]
        for k in KEYS:
            assert mixed[i][k] == want[k], (i, fb[i], k, mixed[i], want)
        assert fb[i] == (1 if (fprops[i]["fallback"] or fprops[i]["confidence"] < thr) else 0), i


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
