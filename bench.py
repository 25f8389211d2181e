"""Throughput benchmark of the CTF-estimation path on one MI355X.

Workload: a batch of synthetic 4096 x 4096 float32 micrographs at 1.0 A
(300 kV, Cs 2.7 mm, amplitude contrast 0.07; defocus, astigmatism, azimuth,
B-factor and SNR vary per micrograph), already resident in HBM.  One step =
one pass over the batch of:
  learned (default)  rescale to 3412^2 at 1.2 A, 144-tile averaged spectrum,
                     background, low-pass, radial average, proposal network,
                     narrowed 2D search (plain search for rejected proposals),
                     spectral-CC resolution
  plain              the same spectrum stages, then the reference's 1D + 2D search
  grid               the spectrum stages, then the 256 x 32 x 36 BF16 grid search

Prints one JSON line.  `value` is the selected path's micrographs/s; in
learned mode the line also carries the plain path's rate on the same batch
and the learned path's defocus agreement with the plain path over a held-out
evaluation set.  Extra objects:
  roofline      the spectrum stages (rescale, tile spectrum, background,
                low-pass) as one HBM-bound unit: their algorithmic bytes per
                micrograph (input read once, rescaled image written and read
                once, four half spectra) times the micrographs of a pass, over
                the summed HIP-event durations of their launches in one extra
                pass on the library's stream; the largest spectrum kernel's
                own figure sits beside it under `per_kernel`
  cpu_baseline  the CPU oracle on a bounded sample of the same workload; its
                results are compared with both GPU paths micrograph by
                micrograph under config.parity_vs_oracle
Further legs in the same run (each skippable): the 256-micrograph BF16 grid
search (config.grid_leg), the uneven list of 64 micrographs
(config.uneven_list), a 41-micrograph tilt series (config.series_leg) and
the GCtfFind command line end to end on a folder of MRC files
(config.files_leg).
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
VALU_PEAK_TFLOPS = 78.6   # FP32 FMA without packed math: 256 CU x 64 lanes x 2 x 2.4 GHz

# HBM bytes per 4096x4096 micrograph at 1.0 A, measured with two rocprofv3
# --pmc passes (FETCH_SIZE, WRITE_SIZE) over `bench.py --mode plain` and
# scaled by the same counters on a 256 MiB device copy (fetch x2.000,
# write x1.000); the table is profiles/pmc_traffic_plain_batch32_4096sq.md.
MEASURED_TRAFFIC_PER_IMAGE = {
    "k_rescale_rows_r2c": 123.08e6, "k_transpose_c": 112.03e6, "k_rescale_cols": 104.02e6,
    "k_rescale_rows_c2r": 93.65e6, "k_tile_rows": 196.52e6, "k_tile_cols": 165.41e6,
}

# Wave64 VALU instructions per 4096x4096 micrograph at 1.0 A (rocprofv3 --pmc
# SQ_INSTS_VALU over `bench.py --mode plain`, all XCDs; the table is
# profiles/pmc_valu_plain_batch32_4096sq.md), and the issue peak they are
# measured against: one wave64 FP32 instruction per CU per cycle (a
# microbenchmark of independent FMA/add/multiply chains reaches 0.97).
MEASURED_VALU_INSTR_PER_IMAGE = {
    "k_rescale_rows_r2c": 9.380e6, "k_rescale_cols": 36.724e6, "k_rescale_rows_c2r": 32.100e6,
}
VALU_ISSUE_PEAK = 256 * 2.4e9          # wave64 instructions per second

# Spectrum-stage kernels (rescale a2, tile spectrum a5-a10, background a11,
# gate a12, low-pass a13) and their measured HBM bytes per micrograph from the
# same PMC table (the small kernels' launches there cover 32 or 16 micrographs).
SPECTRUM_KERNELS = ("k_rescale_rows_r2c", "k_transpose_c", "k_rescale_cols", "k_rescale_rows_c2r",
                    "k_row_sums", "k_block_rows", "k_tile_rows", "k_tile_cols", "k_tile_reduce",
                    "k_bg_rows", "k_bg_cols", "k_gate", "k_lp_x", "k_lp_y")
MEASURED_TRAFFIC_SMALL_PER_IMAGE = {
    "k_block_rows": 0.24e6, "k_tile_reduce": 2.63e6, "k_bg_rows": 1.57e6, "k_bg_cols": 5.29e6,
    "k_gate": 0.01e6, "k_lp_x": 1.61e6, "k_lp_y": 7.21e6,
}


def spectrum_bytes_per_image(nx, ny, in_bytes, nnx, nny, rescaled, tile=512):
    """Algorithmic HBM bytes of the spectrum stages for one micrograph: the
    input read once, the rescaled image written once and read once (when the
    pixel size is below 1.2 A), and four half spectra (raw and
    background-removed, each written and read)."""
    b = nx * ny * in_bytes
    if rescaled:
        b += 2 * nnx * nny * 4
    return b + 4 * (tile // 2 + 1) * tile * 4


# Arithmetic of the FFT kernels per micrograph (5 N log2 N per complex
# transform of length N, 6 per complex multiply, 16 per radix-4 butterfly):
# k_rescale_cols runs 1707 columns of FFT 4096 + an inverse of 3412 = 4 x 853
# as four 853-point Bluestein transforms (each two FFTs of 2048 and three
# pointwise products of 2048) joined by 853 radix-4 butterflies with three
# twiddle products each.
FFT_FLOP_PER_IMAGE = {
    "k_rescale_cols": 1707 * (5 * 4096 * 12 + 4 * (2 * 5 * 2048 * 11 + 3 * 6 * 2048) + 853 * (3 * 6 + 16)),
}


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
    ap.add_argument("--mode", choices=["learned", "plain", "grid"], default="learned",
                    help="learned: proposal network + narrowed search; plain: Do1D + Do2D search; "
                         "grid: 256x32x36 BF16 MFMA grid search")
    ap.add_argument("--eval", type=int, default=1024,
                    help="learned mode: held-out micrographs for the defocus agreement (0: skip)")
    ap.add_argument("--weights", default=None, help="proposal-network weights (.npz)")
    ap.add_argument("--conf-threshold", type=float, default=None,
                    help="learned path: fallback below this confidence (default: the weights' calibrated one)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip gctf_autotune (images per chunk x tile splits chosen by timing)")
    ap.add_argument("--chunk", type=int, default=0, help="images per spectrum chunk (0: library default)")
    ap.add_argument("--splits", type=int, default=0, help="tile-column splits (0: library default)")
    ap.add_argument("--scan1d", default="",
                    help="initial 1D defocus scan lo,hi,step in A for the GPU and the oracle "
                         "(default: the reference's 3000 p^2 .. 30000 p^2 in 500 steps)")
    ap.add_argument("--grid-leg", type=int, default=256, metavar="N",
                    help="also time the BF16 grid search on N 4096^2 micrographs (0: skip); reported "
                         "under config.grid_leg with each result's error against the FP64 oracle")
    ap.add_argument("--series-leg", type=int, default=41, metavar="N",
                    help="also time a tilt series of N 3838x3710 micrographs at 1.35 A (0: skip)")
    ap.add_argument("--uneven-leg", type=int, default=64, metavar="N",
                    help="also time gctf_estimate_list on N uneven micrographs (half 4096^2 f32, a quarter "
                         "5760x4092 f16, a quarter 11520x8184 f16, 0.5-1.1 A); reported under config.uneven_list")
    ap.add_argument("--files-leg", type=int, default=40, metavar="N",
                    help="also time the GCtfFind command line end to end on N 4096^2 float32 MRC files "
                         "(0: skip); reported under config.files_leg")
    ap.add_argument("--no-plain-leg", action="store_true",
                    help="learned mode: do not time the plain path beside it (for profiling)")
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


def make_batch(torch, n, size, pix, seed0=1000, out=None, wide=False):
    """n seeded synthetic micrographs (micrograph i from seed seed0 + i).
    Default: mean defocus log-uniform 8000-30000 A, astigmatism 0-1000 A, B
    100, SNR 0.1.  wide: the proposal network's evaluation distribution --
    defocus log-uniform 3000-50000 A, astigmatism 0-1500 A, B 50-200,
    SNR 0.03-0.3."""
    from gctffind_amd import synth
    import numpy as np
    imgs = out if out is not None else torch.empty((n, size, size), dtype=torch.float32, device="cuda:0")
    truth = []
    from gctffind_amd import proposal
    for i in range(n):
        rng = np.random.default_rng(seed0 + i)
        if wide:
            d1, d2, az, bfac, snr = proposal.wide_params(rng)
            dfm, ast = 0.5 * (d1 + d2), d1 - d2
        else:
            dfm = float(math.exp(rng.uniform(math.log(8000.0), math.log(30000.0))))
            ast = float(rng.uniform(0.0, 1000.0))
            bfac, snr = 100.0, 0.1
            az = float(rng.uniform(-90.0, 90.0))
        synth.micrograph(size, size, pix=pix, seed=seed0 + i, df1=dfm + ast / 2, df2=dfm - ast / 2,
                         az_deg=az, bfac=bfac, snr=snr, device="cuda:0", out=imgs[i])
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
    scan1d = tuple(float(v) for v in a.scan1d.split(",")) if a.scan1d else None
    if scan1d:
        ctx.set_search1d(*scan1d)
    n = a.batch
    learned = a.mode == "learned"
    if learned:
        from gctffind_amd import proposal
        params, meta = proposal.load_params(a.weights or proposal.DEFAULT_WEIGHTS)
        # the confidence threshold calibrated on the tuning seeds with the weights
        thr = float(meta.get("conf_threshold", 0.8)) if a.conf_threshold is None else a.conf_threshold
        ctx.set_proposal_params(params)
    imgs, truth = make_batch(torch, n, a.size, a.pix, seed0=2000 if learned else 1000, wide=learned)
    ptr = imgs.data_ptr()
    tuned = None
    if not a.no_autotune and not a.chunk and not a.splits:
        tuned = ctx.autotune(ptr, n, a.size, a.size, a.pix)

    grid = dict(gctf.GRID_256x32x36)
    plain_run = lambda: ctx.estimate_batch(ptr, n, a.size, a.size, a.pix)
    if a.mode == "grid":
        run = lambda: ctx.grid_search_batch(ptr, n, a.size, a.size, a.pix, grid)
    elif learned:
        run = lambda: ctx.estimate_batch_learned(ptr, n, a.size, a.size, a.pix, conf_threshold=thr)
    else:
        run = plain_run
    log(f"timing the {a.mode} path: batch {n}, {a.steps} steps after {a.warmup} warm-up")
    value, ms, res = timed(torch, run, n, a.steps, a.warmup)

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
            "autotune": tuned,
            "search1d_scan": ({"df_lo_A": scan1d[0], "df_hi_A": scan1d[1], "step_A": scan1d[2],
                               "points": int((scan1d[1] - scan1d[0]) / scan1d[2]) // 2 * 2 + 1}
                              if scan1d else "reference: 3000 p^2 .. 30000 p^2 A in 500 steps "
                                             "(4323-43234 A, 501 points at the working 1.20047 A)"),
        },
    }
    pres = None
    if learned:
        out["config"]["workload"] = (
            f"learned-proposal CTF estimation (rescale + 144-tile spectrum + radial average + proposal "
            f"network + +-300 A narrowed 2D search, plain search below confidence 0.8), batch of {n} "
            f"synthetic {a.size}x{a.size} f32 micrographs at {a.pix} A")
        if not a.no_plain_leg:
            log("timing the plain path on the same batch")
            pv, pms, pres = timed(torch, plain_run, n, a.steps, a.warmup)
            out["config"]["plain_value"] = round(pv, 2)
            out["config"]["plain_ms_per_step"] = round(pms, 3)
            out["config"]["learned_over_plain"] = round(value / pv, 3)
        out["config"]["proposal_weights"] = {k: meta.get(k) for k in ("spectra", "micrograph", "epochs", "steps",
                                                                      "confidence_labels")}
        out["config"]["conf_threshold"] = thr
        if "calibration" in meta:
            out["config"]["conf_calibration"] = {k: meta["calibration"][k] for k in
                                                 ("seeds", "micrograph", "target", "threshold", "table",
                                                  "frac_agree_without_fallback")}
    dfe = [abs(0.5 * (r["df_min"] + r["df_max"]) - t[0]) for r, t in zip(res, truth)]
    out["config"]["median_abs_ddf_vs_truth_A"] = round(float(sorted(dfe)[len(dfe) // 2]), 1)

    if not a.no_roofline:
        log("profiled pass (HIP events per kernel)")
        out["roofline"], out["kernels"] = roofline(ctx, ptr, n, a, run, tuned)
    # the CPU baseline's host copies and the GPU results it is compared with,
    # taken before the evaluation reuses imgs
    host = None
    if not a.no_cpu_baseline:
        host = [imgs[i].cpu().numpy() for i in range(cpu_threads(n))]
        if learned and pres is None:
            pres = ctx.estimate_batch(ptr, n, a.size, a.size, a.pix)
    if learned and a.eval:
        log(f"evaluating {a.eval} held-out micrographs")
        out["config"]["eval"] = evaluate(torch, ctx, a, imgs, thr)
    del imgs
    torch.cuda.empty_cache()
    if a.grid_leg:
        log(f"grid leg: {a.grid_leg} micrographs")
        out["config"]["grid_leg"] = grid_leg(torch, gctf, a)
        torch.cuda.empty_cache()
    if a.uneven_leg:
        log(f"uneven-list leg: {a.uneven_leg} micrographs")
        out["config"]["uneven_list"] = uneven_list(torch, gctf, ctx, a.uneven_leg, a.steps, a.warmup)
        torch.cuda.empty_cache()
    if a.series_leg:
        log(f"tilt-series leg: {a.series_leg} micrographs")
        out["config"]["series_leg"] = series_leg(torch, gctf, a.series_leg, a.steps, a.warmup)
        torch.cuda.empty_cache()
    if a.files_leg:
        log(f"files leg: GCtfFind on {a.files_leg} MRC files")
        out["config"]["files_leg"] = files_leg(torch, a, a.files_leg)
        torch.cuda.empty_cache()
    if not a.no_cpu_baseline:
        log(f"CPU oracle on {len(host)} micrographs")
        gpu = {"plain": pres if learned else res}
        if learned:
            gpu["learned"] = res
        elif a.mode == "grid":
            gpu = {}
        out["cpu_baseline"], out["config"]["parity_vs_oracle"] = cpu_baseline(host, a.pix, gpu, scan1d)
    print(json.dumps(out), flush=True)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed(torch, run, n, steps, warmup):
    """warmup untimed passes, then `steps` timed passes between device syncs."""
    res = None
    for _ in range(warmup):
        res = run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        res = run()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return n * steps / dt, dt * 1000.0 / steps, res


def write_mrc_f32(path, img, pix):
    """One float32 section as an MRC2014 file (little endian, no extended
    header)."""
    import struct
    ny, nx = img.shape
    h = bytearray(1024)
    struct.pack_into("<10i", h, 0, nx, ny, 1, 2, 0, 0, 0, nx, ny, 1)
    struct.pack_into("<6f", h, 40, nx * pix, ny * pix, pix, 90.0, 90.0, 90.0)
    struct.pack_into("<3i", h, 64, 1, 2, 3)
    h[208:212] = b"MAP "
    h[212:214] = b"\x44\x44"
    with open(path, "wb") as f:
        f.write(bytes(h))
        f.write(img.astype("<f4", copy=False).tobytes())


def files_leg(torch, a, n):
    """End to end on files: the GCtfFind command line over a serial folder of
    n MRC files (float32 4096^2 micrographs at 1.0 A, seeds 3000..: one
    collection session, mean defocus uniform 13500-16500 A, astigmatism
    0-1000 A, B 100, SNR 0.1), writing the spectra (-OutMrc) and the CTF
    table (-OutCtf).  The program reads the next files on its reader threads
    while the GPU estimates the current one, and its writer thread saves
    earlier spectra (the CLoadImages / CAsyncSaveImages overlap).  Each
    micrograph goes through the per-micrograph API from host memory; the
    first gets the full search, the others Refine around the best-scoring
    one so far (CFindSeriesCtfs).  value: the micrographs after the first
    over the time the program reports for them ("Steady state"), which
    leaves out its start-up and the first micrograph's plan set-up; the
    whole run's wall time is given beside it.  The files are in the page
    cache (just written)."""
    import re
    import shutil
    import subprocess
    import tempfile
    import numpy as np
    cli = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gctffind_amd", "GCtfFind")
    if not os.path.exists(cli):
        raise RuntimeError("gctffind_amd/GCtfFind is not built (python -m gctffind_amd.build)")
    d = tempfile.mkdtemp(prefix="gctf_files_")
    try:
        from gctffind_amd import synth
        buf = torch.empty((a.size, a.size), dtype=torch.float32, device="cuda:0")
        truth = []
        for i in range(n):
            rng = np.random.default_rng(3000 + i)
            dfm, ast = float(rng.uniform(13500.0, 16500.0)), float(rng.uniform(0.0, 1000.0))
            synth.micrograph(a.size, a.size, pix=a.pix, seed=3000 + i, df1=dfm + ast / 2, df2=dfm - ast / 2,
                             az_deg=float(rng.uniform(-90.0, 90.0)), bfac=100.0, snr=0.1, device="cuda:0", out=buf)
            truth.append(dfm)
            write_mrc_f32(os.path.join(d, f"mic_{i + 1:03d}.mrc"), buf.cpu().numpy(), a.pix)
        del buf
        cmd = [cli, "-InMrc", os.path.join(d, "mic_"), "-InSuffix", ".mrc", "-Serial", "1",
               "-OutMrc", os.path.join(d, "spec.mrc"), "-OutCtf", os.path.join(d, "ctf.txt"), "-PixSize", str(a.pix)]
        t0 = time.perf_counter()
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            raise RuntimeError(f"GCtfFind exited {r.returncode}: {r.stderr[-2000:]}")
        m = re.search(r"Steady state: (\d+) micrographs in ([0-9.]+) \(s\)", r.stdout)
        if not m:
            raise RuntimeError("GCtfFind printed no steady-state time")
        k, ts = int(m.group(1)), float(m.group(2))
        rows = [ln.split() for ln in open(os.path.join(d, "ctf.txt")) if not ln.startswith("#")]
        dfm = np.array([0.5 * (float(x[2]) + float(x[3])) for x in rows])
        err = np.abs(dfm - np.array(truth))
        mb = a.size * a.size * 4 / 1e6
        return {"micrographs": n, "value": round(k / ts, 2), "unit": "micrographs/s",
                "steady_state_s": ts, "steady_state_micrographs": k, "wall_s": round(wall, 3),
                "file": f"{a.size}x{a.size} float32 MRC at {a.pix} A, {mb:.1f} MB",
                "input_MB_per_s": round(k * mb / ts, 1),
                "program": "gctffind_amd/GCtfFind -Serial 1 -OutMrc -OutCtf: first micrograph full search, "
                           "the rest Refine around the best-scoring one",
                "session": "mean defocus uniform 13500-16500 A, astigmatism 0-1000 A",
                "median_abs_ddf_vs_truth_A": round(float(np.median(err)), 1)}
    finally:
        shutil.rmtree(d, ignore_errors=True)


def uneven_list(torch, gctf, ctx, n, steps, warmup):
    """The uneven-batch workload through gctf_estimate_list: n micrographs,
    half 4096^2 float32, a quarter 5760x4092 float16 and a quarter 11520x8184
    float16 (super-resolution), pixel 0.5-1.1 A, defocus log-uniform
    5000-50000 A, astigmatism 0-1500 A, azimuth 0-180 degrees (seeds 200..)."""
    import numpy as np
    from gctffind_amd import synth
    rng = np.random.default_rng(200)
    q = n // 4
    shapes = ([(4096, 4096, torch.float32, gctf.DTYPE_F32)] * (n - 2 * q)
              + [(5760, 4092, torch.float16, gctf.DTYPE_F16)] * q
              + [(11520, 8184, torch.float16, gctf.DTYPE_F16)] * q)
    bufs, images, truth = [], [], []
    for i, (nx, ny, tdt, dt) in enumerate(shapes):
        pix, df1, df2, az = synth.random_params(rng)
        img = synth.micrograph(ny, nx, pix=pix, seed=200 + i, df1=df1, df2=df2, az_deg=az, device="cuda")
        bufs.append(img.to(tdt).contiguous())
        del img
        images.append(dict(data=bufs[-1].data_ptr(), dtype=dt, nx=nx, ny=ny, pix=pix))
        truth.append(0.5 * (df1 + df2))
    torch.cuda.synchronize()
    value, ms, res = timed(torch, lambda: ctx.estimate_list(images), n, steps, warmup)
    dfe = sorted(abs(0.5 * (r["df_min"] + r["df_max"]) - t) for r, t in zip(res, truth))
    mpix = sum(im["nx"] * im["ny"] for im in images) / 1e6
    return {"micrographs": n, "mix": f"{n - 2 * q} x 4096^2 f32, {q} x 5760x4092 f16, {q} x 11520x8184 f16",
            "pixel_A": "0.5-1.1 (uniform)", "value": round(value, 2), "unit": "micrographs/s",
            "ms_per_list": round(ms, 3), "input_mpix_per_list": round(mpix, 1),
            "median_abs_ddf_vs_truth_A": round(float(dfe[len(dfe) // 2]), 1)}


def evaluate(torch, ctx, a, buf, thr):
    """Learned path (confidence threshold thr) against the plain path (the
    reference algorithm; its defocus and azimuth equal the FP64 oracle's on
    the parity cases) on a.eval held-out micrographs,
    seeds 2000..: mean-defocus agreement overall, inside the reference's 1D
    search range (3000 p^2 .. 30000 p^2 at the working pixel size p; outside
    it the reference itself fails) and by defocus band, agreement within the
    plain search's own final defocus step (0.2 df x 0.525 / 31, at least
    20 A), which of the two reached the higher CC, fallback and confidence
    rates, and both paths' error against the generating defocus."""
    import numpy as np
    n = buf.shape[0]
    nnx, nny, pitch, pixn = ctx.rescale_setup(a.size, a.size, a.pix)
    lo1, hi1 = 3000.0 * pixn * pixn, 30000.0 * pixn * pixn
    rows = []
    for i0 in range(0, a.eval, n):
        m = min(n, a.eval - i0)
        _, truth = make_batch(torch, m, a.size, a.pix, seed0=2000 + i0, out=buf, wide=True)
        lr, props = ctx.estimate_batch_learned(buf.data_ptr(), n, a.size, a.size, a.pix, proposals=True,
                                               conf_threshold=thr)
        l0 = ctx.estimate_batch_learned(buf.data_ptr(), n, a.size, a.size, a.pix, conf_threshold=0.0)
        pr = ctx.estimate_batch(buf.data_ptr(), n, a.size, a.size, a.pix)
        for j in range(m):
            rows.append((truth[j][0], 0.5 * (lr[j]["df_min"] + lr[j]["df_max"]), 0.5 * (pr[j]["df_min"] + pr[j]["df_max"]),
                         lr[j]["score"], pr[j]["score"], props[j]["fallback"], props[j]["confidence"],
                         0.5 * (l0[j]["df_min"] + l0[j]["df_max"])))
    t, ml, mp, sl, sp, fb, conf, m0 = (np.array(c, np.float64) for c in zip(*rows))
    dd = np.abs(ml - mp)
    inr = (t >= lo1) & (t <= hi1)
    step = np.maximum(0.2 * mp * 0.525 / 31.0, 20.0)
    lowdf = inr & (t < 24000.0)
    highdf = inr & (t >= 24000.0)
    dis = inr & (dd > 20.0)
    r = lambda x: round(float(x), 4)
    return {"n": int(t.size), "seeds": f"2000..{2000 + t.size - 1}",
            "reference": "plain GPU path (the reference algorithm; defocus and azimuth equal to the FP64 oracle's on the parity cases)",
            "frac_within_20A_all": r(np.mean(dd <= 20.0)),
            "reference_1d_range_A": [round(lo1), round(hi1)],
            "frac_in_reference_range": r(np.mean(inr)),
            "frac_within_20A_in_range": r(np.mean(dd[inr] <= 20.0)),
            "frac_within_20A_in_range_df_below_24000A": r(np.mean(dd[lowdf] <= 20.0)),
            "frac_within_20A_in_range_df_24000A_up": r(np.mean(dd[highdf] <= 20.0)),
            "frac_within_plain_final_step_in_range": r(np.mean(dd[inr] <= step[inr])),
            "frac_learned_cc_ge_plain_where_apart": r(np.mean(sl[dis] >= sp[dis] - 1e-6)) if dis.any() else None,
            "fallback_frac": r(np.mean(fb)),
            "conf_threshold": thr,
            "frac_within_20A_without_confidence_fallback": r(np.mean(np.abs(m0 - mp) <= 20.0)),
            "median_confidence": round(float(np.median(conf)), 3),
            "median_abs_ddf_vs_truth_learned_A": round(float(np.median(np.abs(ml - t))), 1),
            "median_abs_ddf_vs_truth_plain_A": round(float(np.median(np.abs(mp - t))), 1),
            "frac_abs_ddf_vs_truth_over_500A_learned": r(np.mean(np.abs(ml - t) > 500)),
            "frac_abs_ddf_vs_truth_over_500A_plain": r(np.mean(np.abs(mp - t) > 500))}


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
        b = 2 * g["ny"] * g["ncx"] * 8          # the ny x ncx row spectrum read and written transposed
    elif name == "k_rescale_cols":
        b = g["ncx"] * g["ny"] * 8 + g["ncx"] * g["nny"] * 8
    elif name == "k_rescale_rows_c2r":
        b = g["nny"] * g["ncx"] * 8 + g["nny"] * g["pitch"] * 4
    elif name == "k_row_sums":
        b = g["cover"] * 4
    elif name == "k_tile_rows":
        b = g["cover"] * 4 + g["ntiles"] * g["cx"] * g["T"] * 8
    elif name == "k_tile_cols":
        b = g["ntiles"] * g["cx"] * g["T"] * 8 + nsplit * g["cx"] * g["T"] * 4
    else:
        return None
    return b * c


def roofline(ctx, ptr, n, a, run, tuned=None):
    """One extra pass with every kernel bracketed by HIP events on the
    library's stream.  plain / learned: the spectrum stages as one unit --
    their algorithmic bytes (spectrum_bytes_per_image x micrographs of the
    pass) over the summed durations of their launches -- with the largest
    spectrum kernel's own bytes over its average launch beside it.  grid: the
    GEMM's algorithmic flops (2 M K N) over its launch duration."""
    g = geometry(ctx, a.size, a.pix)
    ctx.profile(True)
    run()
    rep = ctx.profile_report()
    ctx.profile(False)
    chunk = min(tuned["chunk"] if tuned else (a.chunk or 4), n)
    nsplit = tuned["tile_splits"] if tuned else (a.splits or 8)
    batch_kernels = {"k_tile_reduce"}
    rows = {}
    for name, (tot, cnt) in rep.items():
        per = chunk if name.startswith(("k_rescale", "k_transpose", "k_block", "k_tile_rows",
                                        "k_tile_cols")) else n
        if name in batch_kernels:
            per = chunk
        avg_ms = tot / cnt
        nb = kernel_bytes(name, g, per, nsplit)
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
    std = a.size == 4096 and a.pix == 1.0
    # the spectrum stages as one unit
    stage = {k: v for k, v in rows.items() if k.split("<")[0] in SPECTRUM_KERNELS}
    stage_ms = sum(v["ms_total"] for v in stage.values())
    b_img = spectrum_bytes_per_image(a.size, a.size, 4, g["nnx"], g["nny"], g["nnx"] != a.size)
    ach = b_img * n / (stage_ms * 1e-3) / 1e9
    traffic = None
    if std:
        per_img = sum(MEASURED_TRAFFIC_PER_IMAGE.get(k.split("<")[0], 0.0)
                      + MEASURED_TRAFFIC_SMALL_PER_IMAGE.get(k.split("<")[0], 0.0) for k in stage)
        traffic = int(per_img * n)
    roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": "spectrum stages: " + " + ".join(sorted(stage)),
            "algorithmic_bytes_per_micrograph": int(b_img), "micrographs_per_pass": n,
            "algorithmic_bytes_per_pass": int(b_img * n), "stage_ms_per_pass": round(stage_ms, 3),
            "stage_us_per_micrograph": round(stage_ms * 1e3 / n, 2),
            "note": "input read once + rescaled image written and read once + 4 half spectra, over the "
                    "summed HIP-event times of the spectrum-stage launches of one pass"}
    # the largest spectrum kernel on its own
    hbm = {k: v for k, v in rows.items() if v["gbs"] is not None}
    dom = max(hbm, key=lambda k: hbm[k]["ms_total"])
    per = chunk
    nb = kernel_bytes(dom, g, per, nsplit)
    pk = {"kernel": dom, "achieved": round(nb / (rows[dom]["avg_us"] * 1e-6) / 1e9, 1), "peak": HBM_PEAK_GBS,
          "unit": "GB/s", "algorithmic_bytes_per_launch": int(nb), "avg_launch_us": rows[dom]["avg_us"],
          "note": "this kernel's compulsory input reads and output writes per launch (the transposed "
                  "spectrum in, the cropped columns out) over its average launch"}
    pk["frac"] = round(pk["achieved"] / HBM_PEAK_GBS, 4)
    base = dom.split("<")[0]
    t1 = MEASURED_TRAFFIC_PER_IMAGE.get(base) if std else None
    pk["traffic"] = None if t1 is None else int(t1 * per)
    fl = FFT_FLOP_PER_IMAGE.get(base) if std else None
    if fl:
        tf = fl * per / (rows[dom]["avg_us"] * 1e-6) / 1e12
        pk["valu"] = {"achieved": round(tf, 1), "peak": VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                      "frac": round(tf / VALU_PEAK_TFLOPS, 4), "flop_per_launch": int(fl * per),
                      "note": "FFT-bound: 5 N log2 N per complex transform of the transforms run (FFT 4096; "
                              "four 853-point Bluestein transforms on 2048 joined by radix-4 butterflies)"}
    vi = MEASURED_VALU_INSTR_PER_IMAGE.get(base) if std else None
    if vi:
        rate = vi * per / (rows[dom]["avg_us"] * 1e-6)
        pk["valu_issue"] = {"achieved": round(rate / 1e9, 1), "peak": round(VALU_ISSUE_PEAK / 1e9, 1),
                            "unit": "G wave64 instructions/s", "frac": round(rate / VALU_ISSUE_PEAK, 4),
                            "instructions_per_launch": int(vi * per),
                            "note": "SQ_INSTS_VALU per launch (PMC) over the live launch time; "
                                    "peak one instruction per CU per cycle at 2.4 GHz"}
    roof["per_kernel"] = pk
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


def cpu_threads(n):
    """Host threads for the CPU baseline: the process's CPUs, at most 16 (a
    one-GPU host's share) and at most n."""
    try:
        cpus = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        cpus = os.cpu_count() or 1
    return max(1, min(16, cpus, n))


def cpu_baseline(imgs, pix, gpu, scan1d=None):
    """The CPU oracle on the host's cores: one micrograph of the same workload
    per thread (the oracle's C calls release the GIL and share no state), each
    with the same 1D scan as the GPU run.  imgs: host micrographs, one per
    thread; gpu: {path name: the GPU's results on the batch}.  Returns the
    baseline and, per GPU path, its per-micrograph error against the oracle."""
    from concurrent.futures import ThreadPoolExecutor
    import numpy as np
    import oracle as O
    O.lib()
    t = len(imgs)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=t) as ex:
        refs = list(ex.map(lambda im: O.estimate(im, pix, scan1d=scan1d)[0], imgs))
    dt = time.perf_counter() - t0
    img = imgs[0]
    par = {"n": t, "oracle": "FP64 oracle, full plain path on the host copy of each micrograph"}
    for name, res in gpu.items():
        g = res[:t]
        d = lambda k: np.array([abs(r[k] - o[k]) for r, o in zip(g, refs)], np.float64)
        dmin, dmax, dcc = d("df_min"), d("df_max"), d("score")
        # the astigmatism azimuth is an axis: 180 degrees apart is the same one
        daz = np.array([abs(math.remainder(r["azimuth"] - o["azimuth"], 180.0)) for r, o in zip(g, refs)])
        dmean = np.array([abs(0.5 * (r["df_min"] + r["df_max"]) - 0.5 * (o["df_min"] + o["df_max"]))
                          for r, o in zip(g, refs)])
        tol = (dmin <= 5.0) & (dmax <= 5.0) & (daz <= 0.5) & (dcc <= 1e-4)
        par[name] = {"max_abs_ddf_min_A": round(float(dmin.max()), 3), "max_abs_ddf_max_A": round(float(dmax.max()), 3),
                     "max_abs_dazimuth_deg": round(float(daz.max()), 4), "max_abs_dcc": float(f"{dcc.max():.3g}"),
                     "frac_mean_df_within_20A": round(float(np.mean(dmean <= 20.0)), 4),
                     "frac_within_all_tolerances": round(float(np.mean(tol)), 4),
                     "abs_dmean_df_A": [round(float(x), 2) for x in dmean]}
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), None)
    except OSError:
        pass
    return ({"value": round(t / dt, 5), "unit": "micrographs/s", "cores": t, "kind": "port",
             "sample": f"{t} micrographs {img.shape[1]}x{img.shape[0]} at {pix} A, one per thread, full plain "
                       f"path (rescale, spectrum, 1D+2D search), FP64 oracle, {dt:.1f} s wall",
             "host": {"cpu": model, "nproc": os.cpu_count()}}, par)


def grid_leg(torch, gctf, a):
    """The BF16 MFMA grid search on a.grid_leg synthetic 4096^2 f32
    micrographs at 1.0 A (256 df x 32 astigmatism x 36 azimuth, ring
    1/30-1/4 1/A): micrographs/s, the GEMM's MFMA rate from HIP events, and
    each result's error against the FP64 oracle evaluating the same
    candidates on the same spectrum (all top-16 re-scores of a sample of
    micrographs), plus the FP64 argmax over the whole grid for one micrograph."""
    from concurrent.futures import ThreadPoolExecutor
    import numpy as np
    import oracle as O
    n, size, pix = a.grid_leg, 4096, 1.0
    grid = dict(gctf.GRID_256x32x36)
    ctx = gctf.Context()
    imgs, truth = make_batch(torch, n, size, pix, seed0=1000)
    ptr = imgs.data_ptr()
    run = lambda: ctx.grid_search_batch(ptr, n, size, size, pix, grid)
    value, ms, _ = timed(torch, run, n, a.steps, a.warmup)
    ctx.profile(True)
    res, cc, top, rs = ctx.grid_search_batch(ptr, n, size, size, pix, grid, debug=True)
    rep = ctx.profile_report()
    ctx.profile(False)
    tot, cnt = rep["k_grid_gemm"]
    gr = ctx_grid_geometry(ctx, a)
    flops = 2.0 * gr["m"] * gr["k"] * n
    ach = flops / (tot / cnt * 1e-3) / 1e12
    del cc
    # FP64 oracle on the same candidates: the GPU's spectrum of micrograph i
    # (single-image path, the batch computes the same spectrum)
    sample = list(range(0, n, max(1, n // 16)))[:16]

    def spectrum(i):
        c1 = gctf.Context()
        nnx, nny, pitch, pixn = c1.rescale_setup(size, size, pix)
        c1.rescale(imgs[i].data_ptr(), on_device=True)
        c1.setup1(300.0, 2.7, 0.07, pixn)
        c1.setup2(nnx, nny)
        c1.gen_half_spectrum(c1.scaled_image())
        sp = c1.get_half_spect(False).astype(np.float64)
        c1.close()
        return sp, pixn
    f = np.float32
    dstep = (f(grid["df_hi"]) - f(grid["df_lo"])) / f(grid["n_df"] - 1)
    astep = (f(grid["ast_hi"]) - f(grid["ast_lo"])) / f(grid["n_ast"] - 1)
    zstep = (f(grid["az_hi"]) - f(grid["az_lo"])) / f(grid["n_az"])

    def cand(m):
        iaz, rest = m % grid["n_az"], m // grid["n_az"]
        iast, idf = rest % grid["n_ast"], rest // grid["n_ast"]
        return f(grid["df_lo"]) + f(idf) * dstep, f(grid["ast_lo"]) + f(iast) * astep, f(grid["az_lo"]) + f(iaz) * zstep

    def one(sp, pixn, m):
        dfm, ast, az = cand(int(m))
        g1 = dict(df_lo=float(dfm), df_hi=float(dfm), n_df=1, ast_lo=float(ast), ast_hi=float(ast), n_ast=1,
                  az_lo=float(az), az_hi=float(az) + 1.0, n_az=1, res_lo=grid["res_lo"], res_hi=grid["res_hi"])
        return O.grid_cc(sp, g1, pixn)[0]
    spects = [spectrum(i) for i in sample]
    errs, chosen_best = [], 0
    with ThreadPoolExecutor(max_workers=cpu_threads(64)) as ex:
        for (sp, pixn), i in zip(spects, sample):
            ref = list(ex.map(lambda m: one(sp, pixn, m), top[i]))
            errs += [abs(float(rs[i, j]) - ref[j]) for j in range(len(ref))]
            chosen_best += int(int(np.argmax(ref)) == int(np.argmax(rs[i])))
        # FP64 over the whole grid for micrograph sample[0], one df row per task
        sp0, pixn0 = spects[0]

        def row(idf):
            dfm = f(grid["df_lo"]) + f(idf) * dstep
            g1 = dict(grid, df_lo=float(dfm), df_hi=float(dfm), n_df=1)
            return O.grid_cc(sp0, g1, pixn0)
        full = np.concatenate(list(ex.map(row, range(grid["n_df"]))))
    m_best = int(np.argmax(full))
    best_in_top = m_best in set(int(v) for v in top[sample[0]])
    gpu_pick = int(top[sample[0]][int(np.argmax(rs[sample[0]]))])
    step = float(dstep)
    dfe = np.array([abs(0.5 * (r["df_min"] + r["df_max"]) - t[0]) for r, t in zip(res, truth)])
    del imgs
    return {"micrographs": n, "value": round(value, 2), "unit": "micrographs/s", "ms_per_batch": round(ms, 3),
            "grid": "256 df (3000-43000 A) x 32 astigmatism (0-0.06) x 36 azimuth, ring 1/30-1/4 1/A",
            "mfma": {"kernel": "k_grid_gemm", "achieved": round(ach, 1), "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(ach / 2500.0, 4), "avg_launch_us": round(tot / cnt * 1e3, 2),
                     "algorithmic_flops_per_launch": flops, "dtype": "bf16 x bf16 -> f32"},
            "vs_fp64_oracle": {"micrographs_checked": len(sample), "candidates_checked": len(errs),
                               "max_abs_dcc_rescored": float(f"{max(errs):.3g}"),
                               "chosen_is_fp64_best_of_its_top16": f"{chosen_best}/{len(sample)}",
                               "fp64_full_grid_best_in_gpu_top16": bool(best_in_top),
                               "gpu_pick_equals_fp64_full_grid_best": gpu_pick == m_best},
            "median_abs_ddf_vs_truth_A": round(float(np.median(dfe)), 1),
            "frac_within_one_df_step_of_truth": round(float(np.mean(dfe <= step)), 4)}


def series_leg(torch, gctf, n, steps, warmup):
    """A tilt series of n 3838x3710 float32 micrographs at 1.35 A, tilts
    -60..60 degrees, defocus 30000 +- 2000 sin(tilt) A at the centre with a
    linear in-plane gradient of pix tan(tilt) A per column, astigmatism
    400 A at 45 degrees, through gctf_estimate_series (the reference's
    package chain on the device, tomo mode)."""
    import numpy as np
    from gctffind_amd import synth
    nx, ny, pix = 3838, 3710, 1.35
    tilts = np.linspace(-60.0, 60.0, n).astype(np.float32)
    imgs = torch.empty((n, ny, nx), dtype=torch.float32, device="cuda:0")
    truth = []
    for i, t in enumerate(tilts):
        dfm = 30000.0 + 2000.0 * math.sin(math.radians(float(t)))
        synth.micrograph(ny, nx, pix=pix, seed=100 + i, df1=dfm + 200.0, df2=dfm - 200.0, az_deg=45.0,
                         device="cuda:0", out=imgs[i], ddf_dx=pix * math.tan(math.radians(float(t))))
        truth.append(dfm)
    torch.cuda.synchronize()
    ctx = gctf.Context()
    value, ms, res = timed(torch, lambda: ctx.estimate_series(imgs.data_ptr(), n, nx, ny, pix, tilts, tomo=True),
                           n, steps, warmup)
    dfe = np.array([abs(0.5 * (r["df_min"] + r["df_max"]) - t) for r, t in zip(res, truth)])
    del imgs
    return {"micrographs": n, "shape": f"{nx}x{ny} f32 at {pix} A", "tilts": "-60..60 degrees",
            "in_plane_gradient": "pix tan(tilt) A per column (defocus span up to 8990 A at 60 degrees)",
            "value": round(value, 2), "unit": "micrographs/s", "ms_per_series": round(ms, 3),
            "median_abs_ddf_vs_truth_A": round(float(np.median(dfe)), 1),
            "max_abs_ddf_vs_truth_A": round(float(dfe.max()), 1)}


if __name__ == "__main__":
    main()
