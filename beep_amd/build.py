"""Builds the HIP shared library libbeep_gpu.so in-tree for gfx950.

The library holds the device kernels, the C ABI of include/beep_gpu.h and the
C++ host model (normalisation and parameter-derived tables)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libbeep_gpu.so")
SOURCES = ["beep_gpu.hip", "host_model.cc", "host/gpu_state.cc"]
HEADERS = ["device_rng.hh", "kernels.hh", "obs_kernels.hh", "step_kernels.hh", "host_model.hh", "host/gpu_state.hh"]
CHECK_SRC = os.path.join(CSRC, "host", "gpu_state_check.cc")
CHECK_BIN = os.path.join(HERE, "bin", "gpu_state_check")


def hipcc():
    for p in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.sep not in p or os.path.exists(p):
            return p
    return "hipcc"


def build(verbose=True):
    """Compiles every source for gfx950 and links libbeep_gpu.so.  Always
    recompiles (the build check must exercise the compiler)."""
    cmd = [hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
           "-I", os.path.join(ROOT, "include"),
           "-o", LIB + ".tmp"] + [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libbeep_gpu.so")
    os.replace(LIB + ".tmp", LIB)
    # The C++ host mirror's self-check program, linked against the library.
    os.makedirs(os.path.dirname(CHECK_BIN), exist_ok=True)
    cmd = [hipcc(), "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(CSRC, "host"),
           CHECK_SRC, "-o", CHECK_BIN + ".tmp", "-L", HERE, "-lbeep_gpu", "-Wl,-rpath,$ORIGIN/.."]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("building gpu_state_check failed")
    os.replace(CHECK_BIN + ".tmp", CHECK_BIN)
    return LIB


if __name__ == "__main__":
    build()
