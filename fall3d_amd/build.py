"""Builds libfall3d_amd.so (in this directory) from csrc/*.hip with hipcc
for gfx950.

The FP64 translation units are compiled with -ffp-contract=off so no
multiply-add is fused and every FP64 rounding follows the reference's
operation order.  The FP32 step units contract multiply-adds within a
source expression only (-ffp-contract=on): with fast contraction the
backend fuses across statements where each kernel's schedule lets it, so
the tile code of the single-context and of the batched kernel rounded
differently.  All HIP code is compiled with packed-FP32 instructions
disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfall3d_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("F3D_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]
EXACT = ["-ffp-contract=off"]
# multiply-adds fused within one expression only, the same in every kernel
# that inlines the tile code
FAST = ["-ffp-contract=on"]

UNITS = {
    "f3d_plain64.hip": EXACT,
    "f3d_meteo.hip": EXACT,
    "f3d_reduce.hip": EXACT,
    "f3d_api.hip": EXACT,
    "f3d_context.hip": EXACT,
    "f3d_fused32.hip": FAST,
    "f3d_flux32.hip": FAST,
    "f3d_batch.hip": EXACT,
    "f3d_noalloc.hip": EXACT,
    "f3d_output.hip": EXACT,
    "f3d_gravity.hip": EXACT,
}


def _fresh(obj, src):
    """The object is newer than its source and every header of csrc/."""
    if not os.path.exists(obj):
        return False
    t = os.path.getmtime(obj)
    deps = [os.path.join(CSRC, src)] + [os.path.join(CSRC, h) for h in os.listdir(CSRC)
                                        if h.endswith(".h")]
    deps.append(os.path.join(HERE, "..", "include", "fall3d_amd.h"))
    return all(os.path.getmtime(d) < t for d in deps)


def _compile(src, flags, incremental=False):
    obj = os.path.join(BUILD, src.replace(".hip", ".o"))
    if incremental and _fresh(obj, src):
        return obj
    cmd = [HIPCC, *COMMON, *flags, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    # the host pass warns about the unknown packed-fp32 feature: harmless
    return obj


def build(verbose=False, incremental=False):
    """Compile every unit (incremental: only those older than their source
    or a header) and link the library and the Fortran driver."""
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(lambda kv: _compile(*kv, incremental=incremental), UNITS.items()))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_driver(verbose)
    return LIB


FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")
DRIVER = os.path.join(HERE, "f3d_driver")


def build_driver(verbose=False):
    """The Fortran host driver (fortran/f3d_driver.f90) linked against the
    library through the ISO_C_BINDING module (fortran/f3d_mod.f90)."""
    if not os.path.exists(FLANG):
        if verbose:
            print("amdflang not found: Fortran driver not built")
        return None
    fdir = os.path.join(HERE, "fortran")
    mod = os.path.join(BUILD, "mod")
    os.makedirs(mod, exist_ok=True)
    cmd = [FLANG, "-O2", "-module-dir", mod, os.path.join(fdir, "f3d_mod.f90"),
           os.path.join(fdir, "f3d_driver.f90"), "-L" + HERE, "-lfall3d_amd",
           "-Wl,-rpath,$ORIGIN", "-o", DRIVER + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"Fortran driver build failed:\n{r.stderr[-4000:]}")
    os.replace(DRIVER + ".tmp", DRIVER)
    if verbose:
        print("built", DRIVER)
    return DRIVER


if __name__ == "__main__":
    build(verbose=True, incremental="--incremental" in sys.argv)
    sys.exit(0)
