"""The host code of the GCtfFind command line -- MRC reading and writing
(mrc_io.cpp), the facade classes (gctf_host.cpp) and the program with its
loader and writer threads (gctffind_main.cpp) -- built with
AddressSanitizer + UndefinedBehaviorSanitizer and, separately, with
ThreadSanitizer, against a CPU stand-in of the C-ABI
(tests/sanitize/gctf_stub.cpp), and run on a tilt-series stack, a serial
folder of float32 / float16 / int16 micrographs, and a folder with a
truncated file.  Every run must end as the real program does and leave no
sanitizer report.  CPU only."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from mrc_util import read_mrc, write_mrc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "gctffind_amd", "csrc", "host")
SOURCES = [os.path.join(HOST, f) for f in ("gctf_host.cpp", "mrc_io.cpp", "gctffind_main.cpp")]
STUB = os.path.join(ROOT, "tests", "sanitize", "gctf_stub.cpp")
FLAGS = {
    "asan": ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"],
    "tsan": ["-fsanitize=thread"],
}
ENV = {
    "asan": {"ASAN_OPTIONS": "detect_leaks=1:halt_on_error=1:abort_on_error=0",
             "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"},
    "tsan": {"TSAN_OPTIONS": "halt_on_error=1:second_deadlock_stack=1"},
}
MARKERS = ("AddressSanitizer", "LeakSanitizer", "runtime error:", "ThreadSanitizer", "UndefinedBehaviorSanitizer")


@pytest.fixture(scope="module")
def binaries(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("sanitize")
    out = {}
    for kind, fl in FLAGS.items():
        exe = str(d / f"GCtfFind_{kind}")
        cmd = [cxx, "-O1", "-g", "-std=c++17", "-pthread", "-fno-omit-frame-pointer", *fl,
               f"-I{os.path.join(ROOT, 'include')}", f"-I{HOST}", *SOURCES, STUB, "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-3000:]
        out[kind] = exe
    return out


def _run(exe, kind, args, timeout=120):
    env = dict(os.environ)
    env.update(ENV[kind])
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=timeout, env=env)
    for m in MARKERS:
        assert m not in r.stderr, (kind, r.stderr[-4000:])
    return r


def _img(seed, ny=600, nx=560):
    rng = np.random.default_rng(seed)
    return (rng.normal(size=(ny, nx)) * 3.0 + 10.0).astype(np.float32)


@pytest.mark.parametrize("kind", ["asan", "tsan"])
def test_tilt_series_stack(binaries, tmp_path, kind):
    stack = np.stack([_img(10 + k) for k in range(5)])
    write_mrc(str(tmp_path / "ts.mrc"), stack, mode=2, tilts=[-6.0, -3.0, 0.0, 3.0, 6.0])
    r = _run(binaries[kind], kind, ["-InMrc", str(tmp_path / "ts.mrc"), "-OutMrc", str(tmp_path / "spec.mrc"),
                                    "-OutCtf", str(tmp_path / "ctf.txt"), "-PixSize", "1.35"])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    rows = [ln for ln in open(tmp_path / "ctf.txt").read().splitlines() if not ln.startswith("#")]
    assert len(rows) == 5
    assert os.path.exists(tmp_path / "ctf_Imod.txt")
    assert read_mrc(str(tmp_path / "spec.mrc")).shape == (5, 512, 512)


@pytest.mark.parametrize("kind", ["asan", "tsan"])
def test_serial_folder(binaries, tmp_path, kind):
    for k, mode in enumerate((2, 12, 1, 2)):
        img = _img(20 + k)
        write_mrc(str(tmp_path / f"mic_{k + 1:03d}.mrc"), img if mode != 1 else np.round(img), mode=mode)
    r = _run(binaries[kind], kind, ["-InMrc", str(tmp_path / "mic_"), "-InSuffix", ".mrc", "-Serial", "1",
                                    "-OutMrc", str(tmp_path / "spec.mrc"), "-OutCtf", str(tmp_path / "ctf.txt"),
                                    "-PixSize", "1.2"])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    rows = [ln for ln in open(tmp_path / "ctf.txt").read().splitlines() if not ln.startswith("#")]
    assert len(rows) == 4
    for k in range(4):
        assert read_mrc(str(tmp_path / f"spec{k + 1:03d}.mrc")).shape == (1, 512, 512)


@pytest.mark.parametrize("kind", ["asan", "tsan"])
def test_serial_folder_truncated_file(binaries, tmp_path, kind):
    for k in range(3):
        write_mrc(str(tmp_path / f"mic_{k + 1:03d}.mrc"), _img(30 + k))
    bad = tmp_path / "mic_002.mrc"
    bad.write_bytes(bad.read_bytes()[:1024 + 4 * 560 * 100])      # header + 100 of 600 rows
    r = _run(binaries[kind], kind, ["-InMrc", str(tmp_path / "mic_"), "-InSuffix", ".mrc", "-Serial", "1",
                                    "-OutMrc", str(tmp_path / "spec.mrc"), "-OutCtf", str(tmp_path / "ctf.txt"),
                                    "-PixSize", "1.2"])
    assert r.returncode == 1, r.stdout[-2000:] + r.stderr[-2000:]
    assert "cannot read" in r.stderr and "mic_002.mrc" in r.stderr
    assert read_mrc(str(tmp_path / "spec001.mrc")).shape == (1, 512, 512)
