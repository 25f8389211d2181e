"""Defocus-proposal network: parameters, training data, training, evaluation.

The network (its kernels are in csrc/gctf_net.hip) reads the first 256 bins
of a micrograph's radial spectrum -- the same background-removed,
low-passed spectrum the 1D search reads (CFindCtf1D.cpp:113-117) -- and
predicts ln(mean defocus / 1e4 A), astigmatism / 1e3 A, azimuth and a
confidence logit.  Training spectra come from this library's own spectrum
stages (gctf_radial_batch) applied to seeded synthetic micrographs.

Usage (on a GPU host):
  python -m gctffind_amd.proposal train --n 1000000 --out gctffind_amd/proposal_net.npz
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import gctf

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "proposal_net.npz")

CH, TAPS, HID, NOUT, FLAT = 64, 7, 128, 4, 1024
COL = CH * TAPS
# (name, weight count, fan-in, bias count) in parameter-vector order
LAYERS = [("conv1", CH * TAPS, TAPS, CH), ("conv2", CH * COL, COL, CH), ("conv3", CH * COL, COL, CH),
          ("conv4", CH * COL, COL, CH), ("fc1", HID * FLAT, FLAT, HID), ("fc2", NOUT * HID, HID, NOUT)]


def n_params():
    return sum(w + b for _, w, _, b in LAYERS)


def layer_slices():
    """{name: (weight slice, bias slice)} of the flat parameter vector."""
    out, o = {}, 0
    for name, w, _, b in LAYERS:
        out[name] = (slice(o, o + w), slice(o + w, o + w + b))
        o += w + b
    return out


def init_params(seed=0):
    """He-normal weights, zero biases; the last layer starts small."""
    rng = np.random.default_rng(seed)
    p = np.zeros(n_params(), np.float32)
    sl = layer_slices()
    for name, w, fan, _ in LAYERS:
        std = math.sqrt(2.0 / fan) * (0.1 if name == "fc2" else 1.0)
        p[sl[name][0]] = rng.normal(0.0, std, w).astype(np.float32)
    # mean-defocus output starts at ln(15000 / 1e4)
    p[sl["fc2"][1].start] = math.log(1.5)
    return p


def save_params(path, params, meta):
    """Parameters (float32) and a JSON description, without pickles."""
    blob = np.frombuffer(json.dumps(meta, sort_keys=True).encode(), dtype=np.uint8)
    np.savez(path, params=np.asarray(params, np.float32), meta=blob)


def load_params(path=DEFAULT_WEIGHTS):
    with np.load(path, allow_pickle=False) as z:
        p = z["params"].astype(np.float32)
        meta = json.loads(bytes(z["meta"]).decode())
    if p.size != n_params():
        raise ValueError(f"{path}: {p.size} parameters, the network has {n_params()}")
    return p, meta


# ---- synthetic training micrographs ----------------------------------------
def micrograph_batch(torch, n, ny, nx, pix, df1, df2, az_deg, bfac, snr, seed, kv=300.0, cs_mm=2.7, ac=0.07,
                     out=None, device="cuda:0"):
    """n float32 micrographs of the image model of gctffind_amd.synth
    (CTF x envelope on a white object, 1 / (1 + (k / 0.02)^2) background of
    the signal's variance, white noise at var(signal) / snr), drawn in one
    step: the sum of the three independent Gaussian fields is a Gaussian
    field whose power spectrum is the sum of theirs,
        P(k) = N (H(k)^2 + c_bg G(k)^2 + var_s / snr),
    with N = nx ny, var_s = sum_k H^2 / N (Parseval) and c_bg = sum H^2 / sum G^2.
    One complex normal draw on the rfft grid and one inverse FFT per image."""
    dev = device
    g = torch.Generator(device=dev)
    g.manual_seed(int(seed))
    ky = torch.fft.fftfreq(ny, d=pix, device=dev)[:, None]
    kx = torch.fft.rfftfreq(nx, d=pix, device=dev)[None, :]
    k2 = kx * kx + ky * ky
    th = torch.atan2(ky.expand_as(k2), kx.expand_as(k2))
    c2t, s2t = torch.cos(2.0 * th), torch.sin(2.0 * th)
    t = lambda v: torch.as_tensor(np.asarray(v, np.float32), device=dev)[:, None, None]
    dmax, dmin = np.maximum(df1, df2), np.minimum(df1, df2)
    mean, half = t(0.5 * (dmax + dmin)), t(0.5 * (dmax - dmin))
    a2 = np.radians(2.0 * np.asarray(az_deg, np.float64))
    df = mean + half * (c2t[None] * t(np.cos(a2)) + s2t[None] * t(np.sin(a2)))
    lam = gctf_wavelength(kv)
    cs = cs_mm * 1e7
    amp = math.atan(ac / math.sqrt(1.0 - ac * ac))
    h2 = torch.sin(math.pi * lam * k2[None] * (df - 0.5 * lam * lam * cs * k2[None]) + amp) ** 2
    h2 *= torch.exp(-t(bfac) * k2[None] / 2.0)
    del df
    g2 = 1.0 / (1.0 + k2 / (0.02 ** 2))
    # full-spectrum sums from the half grid: columns 1 .. nx/2 - 1 stand for two
    wcol = torch.full((k2.shape[1],), 2.0, device=dev)
    wcol[0] = 1.0
    if nx % 2 == 0:
        wcol[-1] = 1.0
    npx = float(nx * ny)
    sh = (h2 * wcol).sum(dim=(1, 2), keepdim=True)
    sg = (g2 * wcol).sum()
    power = h2.add_((sh / sg) * g2[None]).add_((sh / npx) / t(snr))
    amp_k = torch.sqrt(power.mul_(npx / 2.0))
    z = torch.complex(torch.randn(amp_k.shape, generator=g, device=dev),
                      torch.randn(amp_k.shape, generator=g, device=dev)) * amp_k
    del power, amp_k
    img = torch.fft.irfft2(z, s=(ny, nx))
    if out is not None:
        out.copy_(img)
        return out
    return img


def gctf_wavelength(kv):
    return 12.26 / math.sqrt(kv * 1000.0 + 0.9784 * kv * kv)


def draw_params(rng, n, df_lo=3000.0, df_hi=50000.0, ast_max=1500.0, snr=(0.03, 0.3), bfac=(50.0, 200.0)):
    """Per-micrograph (df1, df2, azimuth, B, SNR): defocus log-uniform, the
    astigmatism (df1 - df2) uniform, azimuth uniform in [-90, 90)."""
    dfm = np.exp(rng.uniform(math.log(df_lo), math.log(df_hi), n))
    ast = rng.uniform(0.0, ast_max, n)
    az = rng.uniform(-90.0, 90.0, n)
    b = rng.uniform(bfac[0], bfac[1], n)
    s = np.exp(rng.uniform(math.log(snr[0]), math.log(snr[1]), n))
    return dfm + 0.5 * ast, dfm - 0.5 * ast, az, b, s


AGREE_A = 20.0   # the learned path's acceptance: |mean defocus - plain search's| <= 20 A


def agreement_labels(ctx, ptr, m, size, pix, window=300.0, agree_A=AGREE_A):
    """+1 / -1 per micrograph: whether the learned path without its
    confidence fallback (the network installed in ctx proposes, the
    narrowed search decides; window-edge and score <= 0 fallbacks stay on)
    ends within agree_A of the plain search's mean defocus on the same
    micrograph."""
    plain = ctx.estimate_batch(ptr, m, size, size, pix)
    learned = ctx.estimate_batch_learned(ptr, m, size, size, pix, conf_threshold=0.0, window=window)
    mp = np.array([0.5 * (r["df_min"] + r["df_max"]) for r in plain])
    ml = np.array([0.5 * (r["df_min"] + r["df_max"]) for r in learned])
    return np.where(np.abs(ml - mp) <= agree_A, 1.0, -1.0).astype(np.float32)


def generate(torch, ctx, n_total, size=3328, pix=1.2, batch=16, seed0=10_000_000, log=None, budget_s=None,
             label_every=0, label_params=None):
    """Radial spectra [n][257] (device) and targets [n][4] (device) of
    n_total seeded synthetic micrographs: the batch starting at sample i
    uses seed seed0 + i.  With budget_s, generation stops after the batch
    that crosses the time budget and the first n generated are returned.
    With label_every = k > 0 and label_params (the network parameters that
    propose), every k-th batch is also searched by the plain and the learned
    path, and target[:, 3] holds agreement_labels there (0 elsewhere: no
    label, so those spectra train the proposal but not the confidence)."""
    cx = 257
    radial = torch.empty((n_total, cx), dtype=torch.float32, device="cuda:0")
    target = torch.zeros((n_total, 4), dtype=torch.float32, device="cuda:0")
    imgs = torch.empty((batch, size, size), dtype=torch.float32, device="cuda:0")
    # one staging buffer: the library caches a graph per output address
    stage = torch.empty((batch, cx), dtype=torch.float32, device="cuda:0")
    if label_every:
        ctx.set_proposal_params(label_params)
    t0 = time.time()
    last = t0
    nlab = nagree = 0
    for i0 in range(0, n_total, batch):
        m = min(batch, n_total - i0)
        rng = np.random.default_rng(seed0 + i0)
        d1, d2, az, b, s = draw_params(rng, m)
        micrograph_batch(torch, m, size, size, pix, d1, d2, az, b, s, seed0 + i0, out=imgs[:m])
        torch.cuda.synchronize()
        ctx.radial_batch(imgs.data_ptr(), m, size, size, pix, stage.data_ptr(), cx)
        radial[i0:i0 + m] = stage[:m]
        lab = np.zeros(m, np.float32)
        if label_every and (i0 // batch) % label_every == 0:
            lab = agreement_labels(ctx, imgs.data_ptr(), m, size, pix)
            nlab += m
            nagree += int((lab > 0).sum())
        tg = np.stack([0.5 * (d1 + d2), np.abs(d1 - d2), az, lab], 1).astype(np.float32)
        target[i0:i0 + m] = torch.from_numpy(tg).cuda()
        if log and time.time() - last > 30:
            last = time.time()
            log(f"generated {i0 + m}/{n_total} spectra, {(i0 + m) / (last - t0):.0f}/s"
                + (f", {nlab} labelled ({nagree} agree)" if label_every else ""))
        if budget_s is not None and time.time() - t0 > budget_s:
            torch.cuda.synchronize()
            return radial[:i0 + m], target[:i0 + m]
    torch.cuda.synchronize()
    return radial, target


def train(torch, radial, target, epochs=30, batch=256, lr=1e-3, wd=1e-4, seed=0, params=None, log=None,
          val_frac=0.01):
    """AdamW with linear warm-up and cosine decay; returns (params, report)."""
    n = radial.shape[0]
    nval = max(batch, int(n * val_frac)) if n > 4 * batch else 0
    ntr = n - nval
    net = gctf.Net(batch)
    net.set_params(init_params(seed) if params is None else params)
    steps = epochs * (ntr // batch)
    warm = max(1, steps // 50)
    g = torch.Generator(device="cuda:0")
    g.manual_seed(seed)
    xb = torch.empty((batch, radial.shape[1]), dtype=torch.float32, device="cuda:0")
    tb = torch.empty((batch, 4), dtype=torch.float32, device="cuda:0")
    step, t0, last, losses = 0, time.time(), time.time(), []
    for ep in range(epochs):
        perm = torch.randperm(ntr, generator=g, device="cuda:0")
        for j in range(ntr // batch):
            net.synchronize()   # the trainer's stream is done with xb / tb
            idx = perm[j * batch:(j + 1) * batch]
            torch.index_select(radial, 0, idx, out=xb)
            torch.index_select(target, 0, idx, out=tb)
            torch.cuda.synchronize()
            want = (step % 200 == 0)
            loss, _ = net.loss_grad(xb.data_ptr(), xb.shape[1], tb.data_ptr(), batch, bf16=True, want_loss=want)
            f = (step + 1) / warm if step < warm else 0.5 * (1 + math.cos(math.pi * (step - warm) / max(1, steps - warm)))
            net.adam_step(lr * f, weight_decay=wd)
            if want:
                losses.append((step, loss))
                if log and time.time() - last > 20:
                    last = time.time()
                    log(f"epoch {ep} step {step}/{steps} loss {loss:.4f} ({step / (last - t0):.0f} steps/s)")
            step += 1
    net.synchronize()
    params = net.get_params()
    report = {"steps": steps, "batch": batch, "epochs": epochs, "train_spectra": ntr,
              "train_seconds": round(time.time() - t0, 1), "loss_curve": losses[::max(1, len(losses) // 20)]}
    if nval:
        report["val"] = evaluate_spectra(net, radial[ntr:], target[ntr:], batch)
    net.close()
    return params, report


def decode(out):
    """Raw network outputs -> (mean df A, astigmatism A, azimuth deg, confidence)."""
    o = np.asarray(out, np.float64)
    df = 1e4 * np.exp(np.clip(o[:, 0], -3, 3))
    az = np.degrees(o[:, 2])
    az = az - 180.0 * np.floor((az + 90.0) / 180.0)
    return df, 1e3 * o[:, 1], az, 1.0 / (1.0 + np.exp(-o[:, 3]))


def evaluate_spectra(net, radial, target, batch):
    """Proposal error on held-out spectra."""
    n = radial.shape[0]
    dfs, confs = [], []
    for i0 in range(0, n - batch + 1, batch):
        x = radial[i0:i0 + batch].contiguous()
        df, _, _, conf = decode(net.forward(x.data_ptr(), x.shape[1], batch, bf16=True))
        dfs.append(df)
        confs.append(conf)
    df, conf = np.concatenate(dfs), np.concatenate(confs)
    truth = target[:df.size, 0].cpu().numpy().astype(np.float64)
    err = np.abs(df - truth)
    ok = conf >= 0.8
    return {"n": int(df.size), "median_abs_ddf_A": round(float(np.median(err)), 1),
            "frac_within_300A": round(float(np.mean(err <= 300)), 4),
            "frac_confident": round(float(np.mean(ok)), 4),
            "frac_within_300A_when_confident": round(float(np.mean(err[ok] <= 300)) if ok.any() else 0.0, 4)}


# ---- confidence threshold ----------------------------------------------------
def wide_params(rng):
    """One micrograph of the held-out distribution: mean defocus
    log-uniform 3000-50000 A, astigmatism 0-1500 A, B 50-200 A^2, SNR
    log-uniform 0.03-0.3, azimuth uniform in [-90, 90), drawn in that order.
    Returns (df1, df2, azimuth, B, SNR)."""
    dfm = float(math.exp(rng.uniform(math.log(3000.0), math.log(50000.0))))
    ast = float(rng.uniform(0.0, 1500.0))
    bfac, snr = float(rng.uniform(50.0, 200.0)), float(math.exp(rng.uniform(math.log(0.03), math.log(0.3))))
    az = float(rng.uniform(-90.0, 90.0))
    return dfm + ast / 2, dfm - ast / 2, az, bfac, snr


def wide_batch(torch, n, size, pix, seed0, out):
    """n micrographs of the held-out distribution (micrograph i from seed
    seed0 + i, gctffind_amd.synth's image model) into out[:n]; returns the
    generating mean defocus values."""
    from . import synth
    truth = []
    for i in range(n):
        d1, d2, az, b, snr = wide_params(np.random.default_rng(seed0 + i))
        synth.micrograph(size, size, pix=pix, seed=seed0 + i, df1=d1, df2=d2, az_deg=az, bfac=b, snr=snr,
                         device="cuda:0", out=out[i])
        truth.append(0.5 * (d1 + d2))
    torch.cuda.synchronize()
    return np.array(truth)


THRESHOLDS = (0.0, 0.5, 0.6, 0.7, 0.8, 0.85, 0.9, 0.95, 0.98)


def calibrate(torch, ctx, params, n=1024, seed0=5000, size=4096, pix=1.0, batch=32, target=0.99,
              thresholds=THRESHOLDS, log=None):
    """The confidence threshold on a tuning set (seeds seed0 .., separate
    from the evaluation seeds): per micrograph, the learned path without
    its confidence fallback against the plain search.  A micrograph below
    the threshold takes the plain search's result, so it agrees by
    construction.  Returns the table over `thresholds` (fraction within
    AGREE_A, fallback fraction) and the lowest threshold whose fraction
    reaches `target` (None if none does)."""
    ctx.set_proposal_params(params)
    buf = torch.empty((batch, size, size), dtype=torch.float32, device="cuda:0")
    agree, conf, edge = [], [], []
    for i0 in range(0, n, batch):
        m = min(batch, n - i0)
        wide_batch(torch, m, size, pix, seed0 + i0, buf)
        lr, props = ctx.estimate_batch_learned(buf.data_ptr(), batch, size, size, pix, conf_threshold=0.0,
                                               proposals=True)
        pr = ctx.estimate_batch(buf.data_ptr(), batch, size, size, pix)
        for j in range(m):
            ml = 0.5 * (lr[j]["df_min"] + lr[j]["df_max"])
            mp = 0.5 * (pr[j]["df_min"] + pr[j]["df_max"])
            agree.append(abs(ml - mp) <= AGREE_A)
            conf.append(props[j]["confidence"])
            edge.append(bool(props[j]["fallback"]))
        if log:
            log(f"calibration {i0 + m}/{n}")
    agree, conf, edge = np.array(agree), np.array(conf), np.array(edge)
    table, chosen = [], None
    for t in thresholds:
        fb = edge | (conf < t)
        ok = agree | fb
        row = {"threshold": t, "frac_within_20A": round(float(ok.mean()), 4),
               "fallback_frac": round(float(fb.mean()), 4)}
        table.append(row)
        if chosen is None and row["frac_within_20A"] >= target:
            chosen = t
    return {"n": int(n), "seeds": f"{seed0}..{seed0 + n - 1}", "micrograph": f"{size}x{size} at {pix} A",
            "target": target, "table": table, "threshold": chosen,
            "frac_agree_without_fallback": round(float(agree.mean()), 4)}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train"])
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--size", type=int, default=3328)
    ap.add_argument("--pix", type=float, default=1.2)
    ap.add_argument("--gen-batch", type=int, default=16)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--gen-budget", type=float, default=None, help="seconds of spectrum generation at most")
    ap.add_argument("--label-every", type=int, default=4,
                    help="every k-th generated batch also gets agreement labels for the confidence (0: none)")
    ap.add_argument("--label-weights", default=DEFAULT_WEIGHTS,
                    help="network that proposes while the agreement labels are measured")
    ap.add_argument("--calibrate", type=int, default=1024,
                    help="tuning micrographs (4096^2 at 1.0 A, seeds 5000..) for the confidence threshold (0: skip)")
    a = ap.parse_args(argv)
    import torch
    log = lambda s: print(s, flush=True)
    ctx = gctf.Context()
    t0 = time.time()
    lab_params = load_params(a.label_weights)[0] if a.label_every else None
    radial, target = generate(torch, ctx, a.n, a.size, a.pix, a.gen_batch, log=log, budget_s=a.gen_budget,
                              label_every=a.label_every, label_params=lab_params)
    gen_s = time.time() - t0
    lab = target[:, 3].cpu().numpy()
    log(f"generated {radial.shape[0]} spectra in {gen_s:.0f} s, {int((lab != 0).sum())} with agreement labels")
    params, rep = train(torch, radial, target, a.epochs, a.batch, a.lr, seed=a.seed, log=log)
    meta = {"spectra": int(radial.shape[0]), "micrograph": f"{a.size}x{a.size} at {a.pix} A", "seed0": 10_000_000,
            "generation_seconds": round(gen_s, 1),
            "confidence_labels": {"n": int((lab != 0).sum()), "frac_agree": round(float((lab > 0).sum() / max(1, (lab != 0).sum())), 4),
                                  "rule": f"learned path without the confidence fallback within {AGREE_A:.0f} A of the "
                                          f"plain search, proposals of {os.path.basename(a.label_weights)}"},
            **rep}
    del radial, target
    torch.cuda.empty_cache()
    if a.calibrate:
        cal = calibrate(torch, ctx, params, n=a.calibrate, log=log)
        meta["calibration"] = cal
        meta["conf_threshold"] = cal["threshold"] if cal["threshold"] is not None else 0.8
    save_params(a.out, params, meta)
    log(json.dumps({k: v for k, v in meta.items() if k != "loss_curve"}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
