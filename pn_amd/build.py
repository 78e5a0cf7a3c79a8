"""Builds libpn1.so (HIP, gfx950) and the sticky9 host in-tree.

    python -m pn_amd.build          # or: from pn_amd.build import build; build()

Outputs: pn_amd/libpn1.so and pn_amd/sticky9 (objects in pn_amd/_build/).
Every HIP translation unit is compiled with packed FP32 disabled
(-target-feature -packed-fp32-ops).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libpn1.so")
STICKY9 = os.path.join(HERE, "sticky9")

ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
C_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-Wall", "-Wextra"]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build libpn1)")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build(verbose=False):
    """Compiles every source unconditionally (this is also the build check)."""
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    inc = ["-I", INCLUDE, "-I", CSRC]
    host_o = os.path.join(BUILD, "pn1_host.o")
    api_o = os.path.join(BUILD, "pn1_api.o")
    _run(["gcc", *C_FLAGS, *inc, "-c", os.path.join(CSRC, "pn1_host.c"), "-o", host_o], verbose)
    adj_o = os.path.join(BUILD, "pn1_adj.o")
    _run([hipcc, *HIP_FLAGS, *inc, "-c", os.path.join(CSRC, "pn1_api.hip"), "-o", api_o], verbose)
    # the mixed-form pair kernels: LLVM's max-ILP machine scheduler (see
    # pn1_pair32.hip)
    p32_o = os.path.join(BUILD, "pn1_pair32.o")
    _run([hipcc, *HIP_FLAGS, "-mllvm", "-amdgpu-sched-strategy=max-ilp", *inc, "-c",
          os.path.join(CSRC, "pn1_pair32.hip"), "-o", p32_o], verbose)
    # the adjoint's dual-number kernels: finite values, no signed zeros, so
    # products with zero seed tangents fold away (see pn1_adj.hip)
    _run([hipcc, *HIP_FLAGS, "-ffinite-math-only", "-fno-signed-zeros", *inc, "-c",
          os.path.join(CSRC, "pn1_adj.hip"), "-o", adj_o], verbose)
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", api_o, p32_o, adj_o, host_o, "-o",
          tmp], verbose)
    os.replace(tmp, LIB)
    sticky = os.path.join(CSRC, "sticky9.c")
    if os.path.exists(sticky):
        _run(["gcc", *C_FLAGS, "-I", INCLUDE, sticky, "-o", STICKY9, "-L", HERE, "-lpn1",
              "-Wl,-rpath,$ORIGIN", "-lm"], verbose)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
