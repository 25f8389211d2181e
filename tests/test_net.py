"""Proposal network: the FP64 oracle pinned by finite differences and closed
forms (CPU), parameter layout and weight-file round trip."""
import math

import numpy as np
import pytest

import oracle as O
from gctffind_amd import proposal


def _inputs(n=3, seed=0):
    rng = np.random.default_rng(seed)
    r = rng.normal(size=(n, 257)).astype(np.float32) + np.linspace(3, 0, 257, dtype=np.float32)
    tg = np.stack([np.exp(rng.uniform(np.log(3000), np.log(50000), n)), rng.uniform(0, 1500, n),
                   rng.uniform(-90, 90, n), np.zeros(n)], 1).astype(np.float32)
    return r, tg


def test_layout_matches_oracle():
    assert proposal.n_params() == O.net_param_count() == 218436
    sl = proposal.layer_slices()
    assert sl["fc2"][1].stop == proposal.n_params()


def test_loss_closed_form():
    """All weights zero: every output is its fc2 bias, the loss is closed form."""
    p = np.zeros(proposal.n_params())
    b6 = proposal.layer_slices()["fc2"][1]
    o = np.array([0.4, 0.3, 0.2, -0.7])
    p[b6] = o
    r, tg = _inputs(4)
    tg[:, 3] = [1.0, -1.0, 0.0, 1.0]        # agrees, disagrees, no label, agrees
    L = O.net_loss_grad(p, r, tg, grad=False)
    want = 0.0
    for t in tg.astype(np.float64):
        c, w = (1.0 if t[3] > 0 else 0.0), (1.0 if t[3] != 0 else 0.0)
        pr = 1 / (1 + math.exp(-o[3]))
        want += (100 * (o[0] - math.log(t[0] / 1e4)) ** 2 + 0.25 * (o[1] - t[1] / 1e3) ** 2
                 + 0.05 * (1 - math.cos(2 * (o[2] - math.radians(t[2]))))
                 - 0.5 * w * (c * math.log(pr) + (1 - c) * math.log(1 - pr)))
    assert L == pytest.approx(want / 4, rel=1e-12)
    # d loss / d (confidence bias) = sum of 0.5 w (p - c) / n
    _, g = O.net_loss_grad(p, r, tg)
    pr = 1 / (1 + math.exp(-o[3]))
    assert g[b6.start + 3] == pytest.approx(0.5 * ((pr - 1) + pr + 0.0 + (pr - 1)) / 4, rel=1e-9)
    out = O.net_forward(p, r)
    assert np.array_equal(out, np.tile(o, (4, 1)))


def test_input_normalisation_invariance():
    """The network sees z-scored bins [16, 256): an affine change of the
    radial spectrum (or any change of bins 0..15 and 256) leaves its outputs unchanged."""
    p = proposal.init_params(3).astype(np.float64)
    r, _ = _inputs(2, 5)
    r2 = r * 7.5 - 3.0
    r2[:, :16] = 123.0
    r2[:, 256] = -9.0
    assert np.allclose(O.net_forward(p, r), O.net_forward(p, r2), rtol=1e-6, atol=1e-7)


def test_oracle_gradient_finite_differences():
    """Central differences on every layer's weights and biases."""
    p = proposal.init_params(1).astype(np.float64)
    p[proposal.layer_slices()["fc2"][0]] *= 10.0   # O(1) outputs
    # nonzero biases: with zero biases the zeroed input bins put conv1/conv2
    # pre-activations exactly on the ReLU kink, where differences see half a slope
    rng0 = np.random.default_rng(11)
    for _, bs in proposal.layer_slices().values():
        p[bs] += rng0.normal(0.0, 0.05, bs.stop - bs.start)
    r, tg = _inputs(3, 2)
    tg[:, 3] = [1.0, -1.0, 0.0]
    L, g = O.net_loss_grad(p, r, tg)
    rng = np.random.default_rng(7)
    h = 1e-6
    for name, (ws, bs) in proposal.layer_slices().items():
        for sl in (ws, bs):
            idx = rng.choice(np.arange(sl.start, sl.stop), size=min(4, sl.stop - sl.start), replace=False)
            for i in idx:
                q = p.copy()
                q[i] += h
                lp = O.net_loss_grad(q, r, tg, grad=False)
                q[i] -= 2 * h
                lm = O.net_loss_grad(q, r, tg, grad=False)
                fd = (lp - lm) / (2 * h)
                assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (name, i, fd, g[i])


def test_weight_file_round_trip(tmp_path):
    p = proposal.init_params(9)
    f = tmp_path / "w.npz"
    proposal.save_params(str(f), p, {"spectra": 12, "note": "x"})
    q, meta = proposal.load_params(str(f))
    assert np.array_equal(p, q) and meta["spectra"] == 12


def test_shipped_weights():
    """The trained parameters in the package: the network's size, finite,
    and the training record (spectra from the library's own pipeline)."""
    p, meta = proposal.load_params()
    assert p.size == proposal.n_params() and np.all(np.isfinite(p))
    assert meta["spectra"] >= 500_000 and meta["val"]["median_abs_ddf_A"] < 50.0
