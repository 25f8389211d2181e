"""Throughput benchmark of the CTF-estimation path on one MI355X.

Workload: a batch of synthetic 4096 x 4096 float32 micrographs at 1.0 A
(300 kV, Cs 2.7 mm, amplitude contrast 0.07; defocus, astigmatism and azimuth
vary per micrograph), already resident in HBM.  One step = one pass of the
full plain path over the batch: Fourier-crop rescale to 3412^2 at 1.2 A,
144-tile averaged amplitude spectrum, background removal, low-pass, radial
average, 1D and 2D defocus/astigmatism search, spectral-CC resolution.

Prints one JSON line (see the fields below).  Extra objects:
  roofline      the dominant HBM-bound kernel, timed with HIP events on the
                library's stream during one extra pass
  cpu_baseline  the CPU oracle on a bounded sample of the same workload
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--pix", type=float, default=1.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--mode", choices=["plain", "grid"], default="plain",
                    help="plain: Do1D + Do2D search; grid: 256x32x36 BF16 MFMA grid search")
    ap.add_argument("--chunk", type=int, default=0, help="images per spectrum chunk (0: library default)")
    ap.add_argument("--splits", type=int, default=0, help="tile-column splits (0: library default)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB, --steps times (the known-size "
                         "kernel that scales FETCH_SIZE / WRITE_SIZE counters)")
    return ap.parse_args()


def copy_calibration(torch, mib, steps):
    """A plain copy of known size: each launch reads and writes mib MiB."""
    n = mib * (1 << 20) // 4
    src = torch.rand(n, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "bytes_written": n * 4, "launches": steps}),
          flush=True)


def make_batch(torch, n, size, pix, seed0=1000):
    from gctffind_amd import synth
    import numpy as np
    rng = np.random.default_rng(seed0)
    imgs = torch.empty((n, size, size), dtype=torch.float32, device="cuda:0")
    truth = []
    for i in range(n):
        dfm = float(math.exp(rng.uniform(math.log(8000.0), math.log(30000.0))))
        ast = float(rng.uniform(0.0, 1000.0))
        az = float(rng.uniform(-90.0, 90.0))
        synth.micrograph(size, size, pix=pix, seed=seed0 + i, df1=dfm + ast / 2, df2=dfm - ast / 2,
                         az_deg=az, device="cuda:0", out=imgs[i])
        truth.append((dfm, ast, az))
    torch.cuda.synchronize()
    return imgs, truth


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU: the CTF path has no multi-GPU mode "
              f"(--gpus {a.gpus} requested)", file=sys.stderr)
        sys.exit(2)
    import torch
    if a.copy_calibration:
        copy_calibration(torch, a.copy_calibration, a.steps)
        return
    from gctffind_amd import gctf

    ctx = gctf.Context()
    ctx.tune(chunk=a.chunk, tile_splits=a.splits)
    imgs, truth = make_batch(torch, a.batch, a.size, a.pix)
    n = a.batch
    ptr = imgs.data_ptr()

    grid = dict(gctf.GRID_256x32x36)
    if a.mode == "grid":
        run = lambda: ctx.grid_search_batch(ptr, n, a.size, a.size, a.pix, grid)
    else:
        run = lambda: ctx.estimate_batch(ptr, n, a.size, a.size, a.pix)
    for _ in range(a.warmup):
        res = run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        res = run()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = dt * 1000.0 / a.steps
    value = n * a.steps / dt

    out = {
        "metric": "micrographs/s at 4096², learned-proposal vs plain search, |Δdefocus|≤20 Å",
        "value": round(value, 2),
        "unit": "micrographs/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"plain CTF estimation (rescale + 144-tile spectrum + 1D/2D search), "
                         f"batch of {n} synthetic {a.size}x{a.size} f32 micrographs at {a.pix} A")
            if a.mode == "plain" else
            (f"BF16 MFMA grid search 256 df x 32 astigmatism x 36 azimuth (ring 1/30-1/4 1/A, top-16 "
             f"FP32 re-score) after the spectrum stages, batch of {n} synthetic {a.size}x{a.size} f32 "
             f"micrographs at {a.pix} A"),
            "batch": n, "size": a.size, "pixel_A": a.pix, "tile": 512,
            "path": a.mode,
        },
    }
    dfe = [abs(0.5 * (r["df_min"] + r["df_max"]) - t[0]) for r, t in zip(res, truth)]
    out["config"]["median_abs_ddf_vs_truth_A"] = round(float(sorted(dfe)[len(dfe) // 2]), 1)

    if not a.no_roofline:
        out["roofline"], out["kernels"] = roofline(ctx, ptr, n, a, run)
    if not a.no_cpu_baseline:
        out["cpu_baseline"], par = cpu_baseline(imgs[0].cpu().numpy(), a.pix, res[0])
        out["config"]["parity_vs_oracle_img0"] = par
    print(json.dumps(out), flush=True)


def geometry(ctx, size, pix, tile=512):
    """Shapes of the plain path for a size x size image: the rescaled size from
    the library (CRescaleImage::Setup) and the tile grid of
    CGenAvgSpectrum::SetSizes (integer rules)."""
    nnx, nny, pitch, pixn = ctx.rescale_setup(size, size, pix)
    ov = int(tile * 0.5) // 2 * 2
    step = tile - ov
    ntx, nty = (nnx - ov) // step, (nny - ov) // step
    return dict(nx=size, ny=size, nnx=nnx, nny=nny, pitch=pitch, ncx=nnx // 2 + 1, ntiles=ntx * nty,
                cover=((ntx + 1) * step) * ((nty + 1) * step), cx=tile // 2 + 1, T=tile)


def kernel_bytes(name, g, per_launch_images, nsplit=8):
    """Algorithmic HBM bytes of one launch: the compulsory reads of its input
    and writes of its output for the images it processes."""
    c = per_launch_images
    if name == "k_rescale_rows_r2c":
        b = g["ny"] * g["nx"] * 4 + g["ny"] * g["ncx"] * 8
    elif name == "k_transpose_c":
        b = 0.5 * (2 * g["ny"] * g["ncx"] * 8 + 2 * g["ncx"] * g["nny"] * 8)
    elif name == "k_rescale_cols":
        b = g["ncx"] * g["ny"] * 8 + g["ncx"] * g["nny"] * 8
    elif name == "k_rescale_rows_c2r":
        b = g["nny"] * g["ncx"] * 8 + g["nny"] * g["pitch"] * 4
    elif name == "k_block_sums":
        b = g["cover"] * 4
    elif name == "k_tile_rows":
        b = g["cover"] * 4 + g["ntiles"] * g["cx"] * g["T"] * 8
    elif name == "k_tile_cols":
        b = g["ntiles"] * g["cx"] * g["T"] * 8 + nsplit * g["cx"] * g["T"] * 4
    else:
        return None
    return b * c


def roofline(ctx, ptr, n, a, run):
    """One extra pass with every kernel bracketed by HIP events on the
    library's stream.  plain: the dominant HBM-bound kernel's algorithmic
    bytes per launch over its average launch duration.  grid: the GEMM's
    algorithmic flops (2 M K N) over its launch duration."""
    g = geometry(ctx, a.size, a.pix)
    ctx.profile(True)
    run()
    rep = ctx.profile_report()
    ctx.profile(False)
    chunk = min(a.chunk or 4, n)
    batch_kernels = {"k_tile_reduce"}
    rows = {}
    for name, (tot, cnt) in rep.items():
        per = chunk if name.startswith(("k_rescale", "k_transpose", "k_block", "k_tile_rows",
                                        "k_tile_cols")) else n
        if name in batch_kernels:
            per = chunk
        avg_ms = tot / cnt
        nb = kernel_bytes(name, g, per)
        rows[name] = dict(ms_total=round(tot, 3), launches=cnt, avg_us=round(avg_ms * 1e3, 2),
                          gbs=None if nb is None else round(nb / (avg_ms * 1e-3) / 1e9, 1))
    if a.mode == "grid" and "k_grid_gemm" in rows:
        gr = ctx_grid_geometry(ctx, a)
        flops = 2.0 * gr["m"] * gr["k"] * n
        us = rows["k_grid_gemm"]["avg_us"]
        ach = flops / (us * 1e-6) / 1e12
        roof = {"bound": "mfma", "achieved": round(ach, 1), "peak": 2500.0, "unit": "TFLOP/s",
                "frac": round(ach / 2500.0, 4), "traffic": None, "kernel": "k_grid_gemm",
                "algorithmic_flops_per_launch": flops, "avg_launch_us": us, "dtype": "bf16"}
        return roof, rows
    hbm = {k: v for k, v in rows.items() if v["gbs"] is not None}
    dom = max(hbm, key=lambda k: hbm[k]["ms_total"])
    per = chunk
    nb = kernel_bytes(dom, g, per)
    ach = nb / (rows[dom]["avg_us"] * 1e-6) / 1e9
    roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom,
            "algorithmic_bytes_per_launch": int(nb), "avg_launch_us": rows[dom]["avg_us"]}
    return roof, rows


def ctx_grid_geometry(ctx, a):
    """Candidate count and ring size of the grid search (ring [1/30, 1/4] 1/A
    at the working pixel size, the GCC2D float membership rule)."""
    import numpy as np
    nnx, nny, pitch, pixn = ctx.rescale_setup(a.size, a.size, a.pix)
    f = np.float32
    r2 = f(512) * f(pixn)
    lo, hi = (r2 / f(30.0)) / f(512), (r2 / f(4.0)) / f(512)
    y, x = np.mgrid[0:512, 0:257]
    fy = (y.astype(f) - f(256)) / f(512)
    fx = (f(0.5) * x.astype(f)) / f(256)
    s2 = fx * fx + fy * fy
    k = int(((s2 >= lo * lo) & (s2 <= hi * hi)).sum())
    return {"m": 256 * 32 * 36, "k": k}


def cpu_baseline(img, pix, gpu_res):
    """The CPU oracle (one core) on one micrograph of the same workload."""
    import oracle as O
    t0 = time.perf_counter()
    ref, _ = O.estimate(img, pix)
    dt = time.perf_counter() - t0
    par = {k: round(abs(gpu_res[k] - ref[k]), 4) for k in ("df_min", "df_max", "azimuth", "score")}
    return ({"value": round(1.0 / dt, 5), "unit": "micrographs/s", "cores": 1, "kind": "port",
             "sample": f"1 micrograph {img.shape[1]}x{img.shape[0]} at {pix} A, full plain path "
                       f"(rescale, spectrum, 1D+2D search), FP64 oracle, {dt:.1f} s"}, par)


if __name__ == "__main__":
    main()
