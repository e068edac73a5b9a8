"""In-tree build of libicflow_hip.so (gfx950) and of the CPU oracle used by the tests.

    python -m icflow_amd.build            # library + oracle
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libicflow_hip.so")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # no packed FP32 instructions anywhere (v_pk_{add,mul,fma}_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose, what):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"{what} failed building libicflow_hip.so")


def build_library(verbose=False):
    # the host solve (OpenMP, GNU runtime: the one torch and the oracle load) as an object of
    # its own, linked into the HIP library
    solve_o = os.path.join(HERE, "icf_solve.o")
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-fopenmp", "-c", "-o", solve_o, os.path.join(CSRC, "icf_solve.cpp")],
         verbose, "g++")
    srcs = [os.path.join(CSRC, "icf_lib.hip"), os.path.join(CSRC, "icf_host.cpp")]
    tmp = OUT + ".tmp"
    _run([_hipcc(), *HIP_FLAGS, "-o", tmp, *srcs, "-x", "none", solve_o, "-lgomp"], verbose, "hipcc")
    os.remove(solve_o)
    os.replace(tmp, OUT)
    return OUT


def build_oracle(verbose=False):
    """CPU oracle (test infrastructure): oracle/liboracle.so."""
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-B"], capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("building the oracle failed")
    return os.path.join(ROOT, "oracle", "liboracle.so")


def build(verbose=False):
    build_library(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build(verbose=True)
    print("built", OUT)
