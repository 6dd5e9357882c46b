"""Builds the native libraries in-tree: libopensps_hip.so (the gfx950 kernels and their C ABI, hipcc)
and libopensps.so (the Fortran host of the public C API, flang, linked against libopensps_hip.so)."""
from __future__ import annotations

import glob
import os
import shutil
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIBDIR = os.path.join(PKG, "lib")
HIP_SO = os.path.join(LIBDIR, "libopensps_hip.so")
API_SO = os.path.join(LIBDIR, "libopensps.so")

HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
             # packed FP32 instructions stay off (they have given wrong results on these machines)
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def hipcc() -> str:
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def flang() -> str:
    for c in (shutil.which("flang"), "/opt/rocm/llvm/bin/flang", "/opt/rocm/bin/amdflang"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("flang not found")


def build_hip() -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(PKG, "csrc", "*.hip")))
    tmp = HIP_SO + ".tmp"
    _run([hipcc(), *HIP_FLAGS, *srcs, "-o", tmp])
    os.replace(tmp, HIP_SO)
    return HIP_SO


FORTRAN_ORDER = ["opensps_hip_iface.f90", "opensps_host_types.f90", "opensps_host_filters.f90",
                 "opensps_host_cosmology.f90", "opensps_host_context.f90", "opensps_host_batch.f90",
                 "opensps_host_grad.f90", "opensps_c_api.f90"]


def build_fortran() -> str:
    fdir = os.path.join(PKG, "fortran")
    srcs = [os.path.join(fdir, f) for f in FORTRAN_ORDER if os.path.exists(os.path.join(fdir, f))]
    if not srcs:
        return ""
    mod = os.path.join(LIBDIR, "mod")
    os.makedirs(mod, exist_ok=True)
    tmp = API_SO + ".tmp"
    _run([flang(), "-O2", "-fPIC", "-shared", "-J", mod, *srcs, "-o", tmp, "-L" + LIBDIR, "-lopensps_hip",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, API_SO)
    return API_SO


SHIM_SO = os.path.join(ROOT, "tests", "c", "libbatch_ext_shim.so")


def build_test_shims() -> str:
    """The tests' C shim over the header-only struct entry point (tests/c/batch_ext_shim.c): test infrastructure,
    compiled against include/opensps.h and linked to libopensps.so."""
    src = os.path.join(ROOT, "tests", "c", "batch_ext_shim.c")
    if not os.path.exists(src):
        return ""
    rel = os.path.relpath(LIBDIR, os.path.dirname(SHIM_SO))
    _run(["gcc", "-O2", "-fPIC", "-shared", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"), src, "-o",
          SHIM_SO, "-L" + LIBDIR, "-lopensps", "-Wl,-rpath,$ORIGIN/" + rel])
    return SHIM_SO


def build_all() -> None:
    build_hip()
    build_fortran()


if __name__ == "__main__":
    build_all()
    print("built", HIP_SO, API_SO)
