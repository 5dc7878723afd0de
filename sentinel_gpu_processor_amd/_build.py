"""Build libs1ifg.so (HIP, gfx950) in-tree with hipcc.

Every HIP translation unit is compiled for gfx950 with packed-FP32
instructions disabled, then linked into ``sentinel_gpu_processor_amd/lib``.

Floating-point contraction is expression-local (-ffp-contract=on): a * b + c
inside one expression becomes one FMA, and nothing is fused across
statements.  HIP's default (fuse wherever the optimiser sees a product
feeding a sum) decides per kernel, after inlining, so one device function
could round differently in the per-pair and the batched kernels; with
expression-local contraction every kernel that inlines it computes the same
bits.
"""

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
LIBNAME = "libs1ifg.so"
ARCH = os.environ.get("S1IFG_ARCH", "gfx950")

SOURCES = ["host.cpp", "k_geom.hip", "k_ifg.hip", "k_goldstein.hip", "k_gold_mfma.hip",
           "k_batch.hip", "k_adjoint.hip", "k_stitch.hip", "k_ps.hip"]

HIPFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=on",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def lib_path():
    return os.path.join(LIBDIR, LIBNAME)


def build(verbose=False, jobs=None):
    hipcc = _hipcc()
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc] + HIPFLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        if verbose:
            msg = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
            if msg.strip():
                print(msg, file=sys.stderr)
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs or min(8, len(SOURCES))) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    out = lib_path()
    tmp = out + ".tmp"
    cmd = [hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
