"""Proposal network on the GPU against the FP64 oracle: forward pass and
gradients through the C-ABI trainer (FP32-operand MFMA to 1e-4, BF16-operand
MFMA to the BF16 tolerance), AdamW against its formula, a short training run,
the radial-spectrum batch against the single-image spectrum, and the learned
path's fallback against the plain path."""
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


def _params(seed):
    from gctffind_amd import proposal
    p = proposal.init_params(seed)
    sl = proposal.layer_slices()
    p[sl["fc2"][0]] *= 10.0
    rng = np.random.default_rng(seed + 100)
    for _, bs in sl.values():
        p[bs] += rng.normal(0.0, 0.05, bs.stop - bs.start).astype(np.float32)
    return p


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    x = np.linspace(0, 1, 257, dtype=np.float32)
    r = (rng.normal(size=(n, 257)) * 0.5 + np.sin(40 * x[None] ** 2 * rng.uniform(0.5, 2, (n, 1)))).astype(np.float32)
    # confidence labels of every kind: agrees (+1), disagrees (-1), unlabelled (0)
    tg = np.stack([np.exp(rng.uniform(np.log(3000), np.log(50000), n)), rng.uniform(0, 1500, n),
                   rng.uniform(-90, 90, n), rng.choice([-1.0, 0.0, 1.0], n)], 1).astype(np.float32)
    return r, tg


def test_forward_matches_oracle():
    gctf = _setup()
    n = 37
    p = _params(0)
    r, _ = _inputs(n, 1)
    net = gctf.Net(64)
    net.set_params(p)
    d = torch.from_numpy(r).cuda()
    torch.cuda.synchronize()
    ref = O.net_forward(p.astype(np.float64), r)
    scale = max(1.0, float(np.abs(ref).max()))
    out32 = net.forward(d.data_ptr(), 257, n, bf16=False)
    assert np.abs(out32 - ref).max() <= 1e-4 * scale
    out16 = net.forward(d.data_ptr(), 257, n, bf16=True)
    assert np.abs(out16 - ref).max() <= 3e-2 * scale
    net.close()


def test_gradients_match_oracle():
    """Per-layer relative gradient error: FP32-operand MFMA within 1e-4 of
    FP64; BF16 operands within 1e-1 (weight gradients sum terms of both signs
    over batch x length, so the 2^-9 operand rounding grows) and direction
    cosine above 0.995.  n = 37
    exercises ragged GEMM tiles and the split-K reduction (K = 37 x 128)."""
    gctf = _setup()
    from gctffind_amd import proposal
    n = 37
    p = _params(2)
    r, tg = _inputs(n, 3)
    net = gctf.Net(64)
    net.set_params(p)
    dr, dt = torch.from_numpy(r).cuda(), torch.from_numpy(tg).cuda()
    torch.cuda.synchronize()
    lref, gref = O.net_loss_grad(p.astype(np.float64), r, tg)
    for bf16, tol in ((False, 1e-4), (True, 1e-1)):
        loss, g = net.loss_grad(dr.data_ptr(), 257, dt.data_ptr(), n, bf16=bf16, want_grad=True)
        assert loss == pytest.approx(lref, rel=1e-5 if not bf16 else 2e-2)
        for name, (ws, bs) in proposal.layer_slices().items():
            for sl in (ws, bs):
                a, b = g[sl].astype(np.float64), gref[sl]
                nb = np.linalg.norm(b)
                if nb == 0:
                    assert np.abs(a).max() == 0
                    continue
                rel = np.linalg.norm(a - b) / nb
                cos = a @ b / (np.linalg.norm(a) * nb)
                assert rel <= tol, (bf16, name, sl, rel)
                if bf16:
                    assert cos >= 0.995, (name, sl, cos)
    net.close()


def test_adam_step_matches_formula():
    gctf = _setup()
    n = 16
    p = _params(4)
    r, tg = _inputs(n, 5)
    net = gctf.Net(n)
    net.set_params(p)
    dr, dt = torch.from_numpy(r).cuda(), torch.from_numpy(tg).cuda()
    torch.cuda.synchronize()
    m = np.zeros(p.size)
    v = np.zeros(p.size)
    q = p.astype(np.float64)
    lr, b1, b2, eps, wd = 1e-3, 0.9, 0.999, 1e-8, 1e-2
    for t in (1, 2):
        _, g = net.loss_grad(dr.data_ptr(), 257, dt.data_ptr(), n, bf16=False, want_grad=True)
        net.adam_step(lr, b1, b2, eps, wd)
        g = g.astype(np.float64)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        q = q - lr * ((m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps) + wd * q)
        got = net.get_params().astype(np.float64)
        assert np.abs(got - q).max() <= 1e-6
        net.set_params(got.astype(np.float32))   # same starting point for the next comparison
        q = got
    net.close()


def test_radial_batch_matches_single_image_spectrum():
    gctf = _setup()
    from gctffind_amd import synth
    pix = 1.2
    imgs = torch.empty((3, 1024, 1024), dtype=torch.float32, device="cuda")
    for i in range(3):
        synth.micrograph(1024, 1024, pix=pix, seed=70 + i, df1=9000 + 4000 * i, df2=8800 + 4000 * i,
                         az_deg=10.0, device="cuda", out=imgs[i])
    out = torch.zeros((3, 300), dtype=torch.float32, device="cuda")
    ctx = gctf.Context()
    ctx.radial_batch(imgs.data_ptr(), 3, 1024, 1024, pix, out.data_ptr(), 300)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    for i in range(3):
        c1 = gctf.Context()
        nnx, nny, pitch, pixn = c1.rescale_setup(1024, 1024, pix)
        c1.rescale(imgs[i].data_ptr(), on_device=True)
        c1.setup1(300.0, 2.7, 0.07, pixn)
        c1.setup2(nnx, nny)
        c1.gen_half_spectrum(c1.scaled_image())
        sp = c1.get_half_spect(False).astype(np.float64)
        ref = O.radial_avg(sp)
        assert np.allclose(got[i, :257], ref, rtol=1e-6, atol=1e-7 * np.abs(ref).max())
        assert np.all(got[i, 257:] == 0)


def test_training_reduces_loss():
    """Spectra from the library's own pipeline, 300 AdamW steps on BF16 MFMA."""
    gctf = _setup()
    from gctffind_amd import proposal
    ctx = gctf.Context()
    radial, target = proposal.generate(torch, ctx, 768, size=1024, pix=1.2, batch=16, seed0=123)
    params, rep = proposal.train(torch, radial, target, epochs=25, batch=64, lr=2e-3, val_frac=0.0)
    curve = rep["loss_curve"]
    assert len(curve) >= 2 and curve[-1][1] < 0.7 * curve[0][1], curve
    assert np.all(np.isfinite(params))


def test_learned_path_fallback_equals_plain():
    """With a confidence threshold no proposal reaches, every micrograph
    falls back to the plain search: results equal gctf_estimate_batch's."""
    gctf = _setup()
    from gctffind_amd import proposal, synth
    pix = 1.2
    n = 4
    imgs = torch.empty((n, 1024, 1024), dtype=torch.float32, device="cuda")
    for i in range(n):
        synth.micrograph(1024, 1024, pix=pix, seed=80 + i, df1=12000 + 3000 * i, df2=11500 + 3000 * i,
                         az_deg=-20.0 + 15 * i, device="cuda", out=imgs[i])
    ctx = gctf.Context()
    plain = ctx.estimate_batch(imgs.data_ptr(), n, 1024, 1024, pix)
    ctx.set_proposal_params(proposal.init_params(0))
    res, props = ctx.estimate_batch_learned(imgs.data_ptr(), n, 1024, 1024, pix, conf_threshold=2.0,
                                            proposals=True)
    assert all(p["fallback"] == 1 for p in props)
    for a, b in zip(res, plain):
        for k in ("df_min", "df_max", "azimuth", "score", "ctf_res"):
            assert a[k] == b[k], (k, a, b)
