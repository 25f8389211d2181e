"""GCtfFind command line on a small tilt series against the oracle driving
the reference's package logic (zero tilt first with a full search, then
Refine around the best-scoring image so far; CFindSeriesCtfs.cpp:133-203)."""
import math
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from mrc_util import read_mrc, write_mrc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gctffind_amd", "GCtfFind")


def _gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def oracle_series(imgs, pix, tilts, tomo=True):
    """The reference's package loop restated with the oracle: a tilt series
    starts at zero tilt and keeps astigmatism and azimuth fixed in Refine;
    other inputs go in file order and refine them within 0.04 and 30 degrees
    (CFindSeriesCtfs.cpp:168-203)."""
    n = len(imgs)
    z = int(np.argmin(np.abs(tilts))) if tomo else 0
    order = [z] + [i for i in range(n) if i != z]
    res = [None] * n
    ref = None
    for done, i in enumerate(order):
        nnx, nny, pitch, pixn, _ = O.rescale_sizes(imgs[i].shape[1], imgs[i].shape[0], pix)
        fc = O.FindCtf2D(pix=pixn)
        fc.gen_half_spectrum(O.rescale(imgs[i], pix), nnx, nny)

        def full():
            fc.set_phase(0.0, 0.0)
            return fc.do2d()
        if done == 0:
            r = full()
            ref = i
        else:
            rp = res[ref]
            dm = 0.5 * (rp["df_min"] + rp["df_max"])
            dfr = [dm, max(dm * 0.5, 5000.0 * pix * pix)]
            mean = (rp["df_max"] + rp["df_min"]) * 0.5
            ast = [0.0 if mean <= 0 else (rp["df_max"] - rp["df_min"]) * 0.5 / mean, 0.0]
            if not tomo:
                ast[1] = 0.04
            r = fc.refine(np.float32(dfr), np.float32(ast), [rp["azimuth"], 0.0 if tomo else 30.0],
                          [rp["ext_phase"], 0.0])
            if res[ref]["score"] < r["score"]:
                ref = i
            if r["score"] <= 0:
                r = full()
                ref = i
        res[i] = r
    return res


def test_tilt_series_cli_matches_oracle(tmp_path):
    _gpu()
    from gctffind_amd import synth
    pix = 1.35
    tilts = np.array([-6.0, -3.0, 0.0, 3.0, 6.0], np.float32)
    imgs = []
    for k, t in enumerate(tilts):
        df = 30000.0 + 2000.0 * math.sin(math.radians(t))
        imgs.append(synth.micrograph_np(1024, 1024, pix=pix, seed=300 + k, df1=df + 200.0, df2=df - 200.0,
                                        az_deg=45.0))
    stack = str(tmp_path / "ts.mrc")
    write_mrc(stack, np.stack(imgs))
    ang = tmp_path / "ts.tlt"
    ang.write_text("".join(f"{t:.2f}\n" for t in tilts))
    out_ctf, out_mrc = str(tmp_path / "ts_ctf.txt"), str(tmp_path / "ts_spec.mrc")
    r = subprocess.run([CLI, "-InMrc", stack, "-OutMrc", out_mrc, "-OutCtf", out_ctf, "-AngFile", str(ang),
                        "-PixSize", str(pix), "-kV", "300", "-Cs", "2.7", "-AmpContrast", "0.07"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    rows = [ln.split() for ln in open(out_ctf).read().splitlines() if not ln.startswith("#")]
    assert len(rows) == 5
    ref = oracle_series(imgs, pix, tilts)
    for i, row in enumerate(rows):
        assert int(row[0]) == i and float(row[1]) == pytest.approx(tilts[i], abs=0.01)
        dfmin, dfmax, az, ph, score = map(float, row[2:7])
        assert abs(dfmin - ref[i]["df_min"]) <= 5.0, (i, row, ref[i])
        assert abs(dfmax - ref[i]["df_max"]) <= 5.0, (i, row, ref[i])
        assert abs(az - ref[i]["azimuth"]) <= 0.5
        assert abs(score - ref[i]["score"]) <= 2e-4        # printed with %8.4f
    imod = open(out_ctf.replace(".txt", "_Imod.txt")).read().splitlines()
    assert imod[0] == "1  0  0.0  0.0  0.0  3" and len(imod) == 6
    spec = read_mrc(out_mrc)
    assert spec.shape == (5, 512, 512) and np.isfinite(spec).all()


@pytest.mark.parametrize("pix", [1.35, 0.9])
def test_serial_folder_cli(tmp_path, pix):
    """-Serial 1: every micrograph of a folder is estimated (read ahead on the
    loader thread, float32 and float16 files) and its full spectrum written to
    a file of its own (on the writer thread).  Each micrograph after the first
    reaches the GPU through the copy started while the one before it was
    estimated (gctf_rescale_prefetch): at 1.35 A into the unrescaled image,
    at 0.9 A into the Fourier-crop rescale."""
    _gpu()
    from gctffind_amd import synth
    truth = [15000.0, 16500.0, 18000.0]
    for k, df in enumerate(truth):
        img = synth.micrograph_np(1024, 1024, pix=pix, seed=320 + k, df1=df + 250.0, df2=df - 250.0,
                                  az_deg=30.0 * k)
        write_mrc(str(tmp_path / f"mic_{k + 1:03d}.mrc"), img, mode=12 if k == 1 else 2)
    out_ctf, out_mrc = str(tmp_path / "ctf.txt"), str(tmp_path / "spec.mrc")
    r = subprocess.run([CLI, "-InMrc", str(tmp_path / "mic_"), "-InSuffix", ".mrc", "-Serial", "1",
                        "-OutMrc", out_mrc, "-OutCtf", out_ctf, "-PixSize", str(pix)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    rows = [ln.split() for ln in open(out_ctf).read().splitlines() if not ln.startswith("#")]
    assert len(rows) == 3
    imgs = [read_mrc(str(tmp_path / f"mic_{k + 1:03d}.mrc"))[0].astype(np.float32) for k in range(3)]
    ref = oracle_series(imgs, pix, np.zeros(3, np.float32), tomo=False)
    for i, row in enumerate(rows):
        dfmin, dfmax, az = map(float, row[2:5])
        assert abs(dfmin - ref[i]["df_min"]) <= 5.0 and abs(dfmax - ref[i]["df_max"]) <= 5.0, (i, row, ref[i])
        assert abs(az - ref[i]["azimuth"]) <= 0.5, (i, row, ref[i])
        spec = read_mrc(str(tmp_path / f"spec{i + 1:03d}.mrc"))
        assert spec.shape == (1, 512, 512) and np.isfinite(spec).all()


def test_serial_folder_cli_truncated_file(tmp_path):
    """A truncated micrograph in a serial folder ends the run with a read
    error (exit status 1) after the micrographs before it; the loader and
    writer threads are joined, so the program exits instead of hanging."""
    _gpu()
    from gctffind_amd import synth
    pix = 1.35
    for k in range(3):
        img = synth.micrograph_np(1024, 1024, pix=pix, seed=330 + k, df1=15250.0, df2=14750.0, az_deg=20.0)
        write_mrc(str(tmp_path / f"mic_{k + 1:03d}.mrc"), img)
    bad = tmp_path / "mic_002.mrc"
    bad.write_bytes(bad.read_bytes()[:1024 + 4 * 1024 * 100])     # header + 100 of 1024 rows
    r = subprocess.run([CLI, "-InMrc", str(tmp_path / "mic_"), "-InSuffix", ".mrc", "-Serial", "1",
                        "-OutMrc", str(tmp_path / "spec.mrc"), "-OutCtf", str(tmp_path / "ctf.txt"),
                        "-PixSize", str(pix)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1, r.stdout[-2000:] + r.stderr[-2000:]
    assert "cannot read" in r.stderr and "mic_002.mrc" in r.stderr
    assert re.search(r"mic_001\.mrc:\s+1 prcoessed", r.stdout), r.stdout[-2000:]   # the first was processed
    spec = read_mrc(str(tmp_path / "spec001.mrc"))   # and its spectrum written
    assert spec.shape == (1, 512, 512) and np.isfinite(spec).all()
