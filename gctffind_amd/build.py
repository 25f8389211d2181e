"""Build libgctf_hip.so (gfx950) in-tree with hipcc.

Usage: python -m gctffind_amd.build [--debug]
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgctf_hip.so")
ARCH = os.environ.get("GCTF_OFFLOAD_ARCH", "gfx950")

SOURCES = ["gctf_engine.hip", "gctf_net.hip"]
HEADERS = ["gctf_fft.hpp", "gctf_internal.hpp", "k_spectrum.hpp", "k_search.hpp"]


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found")
    return exe


def flags(debug=False):
    f = [
        f"--offload-arch={ARCH}",
        "-std=c++17",
        "-fPIC",
        "-shared",
        # every float expression is evaluated as written (no fused contractions);
        # kernels write fmaf explicitly where they want it
        "-ffp-contract=off",
        # packed FP32 instructions are kept out of the generated code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-fvisibility=hidden",
        "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    ]
    f += ["-O0", "-g"] if debug else ["-O3"]
    return f


HOST_SOURCES = ["host/gctf_host.cpp", "host/mrc_io.cpp", "host/gctffind_main.cpp"]
CLI = os.path.join(HERE, "GCtfFind")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")


def build_cli(verbose=True):
    """The GCtfFind command line (C++ over the C-ABI), linked to libgctf_hip.so."""
    cxx = shutil.which("g++") or "g++"
    srcs = [os.path.join(CSRC, s) for s in HOST_SOURCES]
    tmp = CLI + ".tmp"
    cmd = [cxx, "-O2", "-std=c++17", "-pthread", "-Wall", "-ffp-contract=off", f"-I{INCLUDE}",
           f"-I{os.path.join(CSRC, 'host')}", "-o", tmp] + srcs + [
        f"-L{HERE}", "-lgctf_hip", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib",
        "-L/opt/rocm/lib", "-lamdhip64"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, CLI)
    return CLI


def build(debug=False, verbose=True):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + flags(debug) + ["-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    build_cli(verbose)
    return LIB


if __name__ == "__main__":
    build(debug="--debug" in sys.argv)
