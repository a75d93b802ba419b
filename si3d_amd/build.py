"""Build libsi3dgpu.so for gfx950 with hipcc (in-tree, no JIT cache).

Each translation unit is compiled in parallel, then linked into
si3d_amd/_build/libsi3dgpu.so.  Packed-FP32 instructions are disabled for
all device code.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libsi3dgpu.so")
SOURCES = ["s3g_kernels.hip", "s3g_pcg.hip", "s3g_host.hip", "s3g_unet.hip", "s3g_deck.hip", "s3g_openbc.hip"]
ARCH = os.environ.get("SI3D_OFFLOAD_ARCH", "gfx950")
# The hydrodynamic kernels are built without fused multiply-add contraction:
# each has several instantiations (per-cell chunked or not, split or fused
# phases) that the autotuner picks between, and they must round alike so
# that results do not depend on the pick.
NO_CONTRACT = {"s3g_kernels.hip", "s3g_openbc.hip"}


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src):
    obj = os.path.join(OUT, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), "-c", "-fPIC", "-O3", "-std=c++17", f"--offload-arch={ARCH}",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-result", "-I", os.path.join(HERE, "..", "include"),
           os.path.join(CSRC, src), "-o", obj]
    if src in NO_CONTRACT:
        cmd.insert(5, "-ffp-contract=off")
    if src.endswith(".cpp"):
        cmd = [hipcc(), "-c", "-fPIC", "-O3", "-std=c++17", "-x", "c++",
               "-I", os.path.join(HERE, "..", "include"), os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{r.stdout}\n{r.stderr}")
    return obj


def build_fortran_binding(verbose=False):
    """Compile the ISO_C_BINDING module with flang when the image has it (a
    check that the Fortran side of the boundary stays consistent)."""
    flang = None
    for c in (os.environ.get("FLANG"), "/opt/rocm/lib/llvm/bin/flang", shutil.which("flang")):
        if c and os.path.exists(c):
            flang = c
            break
    if flang is None:
        return None
    out = os.path.join(OUT, "fortran")
    os.makedirs(out, exist_ok=True)
    src = os.path.join(HERE, "fortran", "si3d_gpu_mod.f90")
    obj = os.path.join(out, "si3d_gpu_mod.o")
    r = subprocess.run([flang, "-c", "-fPIC", src, "-o", obj], cwd=out, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"flang failed on si3d_gpu_mod.f90\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built", obj)
    # a Fortran caller of the library through the module, as a shared library
    # (the Fortran binding's check: tests call it in-process)
    chk = os.path.join(HERE, "fortran", "s3g_fortran_check.f90")
    lib = os.path.join(out, "libs3g_fortran_check.so")
    r = subprocess.run([flang, "-shared", "-fPIC", chk, obj, "-L", OUT, "-lsi3dgpu", "-Wl,-rpath,$ORIGIN/..",
                        "-o", lib], cwd=out, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"flang failed on s3g_fortran_check.f90\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built", lib)
    return obj


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(4, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_fortran_binding(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
