"""Build libocean_gpu.so (gfx950) in-tree with hipcc, and the Fortran host pieces."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libocean_gpu.so")
SRC = [os.path.join(HERE, "csrc", "api.hip")]
DEPS = [os.path.join(HERE, "csrc", f) for f in ("common.hpp", "action.hpp", "vec.hpp", "obf.hpp", "gemm64.hpp", "setup.hpp")] + \
    [os.path.join(ROOT, "include", "ocean_gpu.h")]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # keep hipcc from forming packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc") or ""):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build_lib(verbose=True):
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I" + os.path.join(ROOT, "include")] + SRC + ["-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def build_fortran(verbose=True):
    """Compile the iso_c_binding module and the Fortran host driver (amdflang)."""
    fdir = os.path.join(HERE, "fortran")
    fc = shutil.which("amdflang") or shutil.which("flang")
    if fc is None or not os.path.isdir(fdir):
        return None
    out = os.path.join(fdir, "ocean_gpu_driver.x")
    srcs = [os.path.join(fdir, "ocean_gpu_iface.f90"), os.path.join(fdir, "ocean_gpu_host.f90"),
            os.path.join(fdir, "ocean_gpu_driver.f90")]
    srcs = [s for s in srcs if os.path.exists(s)]
    if len(srcs) < 3:
        return None
    link = ["-L" + HERE, "-locean_gpu", "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath," + HERE]
    cmd = [fc, "-O2", "-J", fdir] + srcs + ["-o", out] + link
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    # o_spect.x over the same C-ABI
    osrc = os.path.join(fdir, "ocean_gpu_ospect.f90")
    if os.path.exists(osrc):
        cmd = [fc, "-O2", "-J", fdir, srcs[0], osrc, "-o", os.path.join(fdir, "ocean_gpu_ospect.x")] + link
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    return out


def build(verbose=True):
    build_lib(verbose)
    build_fortran(verbose)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
