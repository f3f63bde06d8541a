"""Compile the HIP library ``pdac_amd/libpdacgpu.so`` for gfx950.

    python -m pdac_amd.build

The kernels keep the reference's operand order, so the build disables
floating-point contraction and fast-math; packed FP32 instructions are
disabled as well (the kernels are FP64)."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpdacgpu.so")
SOURCES = ["pdac_kernels.hip"]
DEPS = ["pdac_kernels.hip", "pdac_iter.hip.inc", "pdac_energy.hip.inc", "pdac_explicit.hip.inc",
        "pdac_ensemble.hip.inc", "pdac_api.hip.inc",
        "pdac_dev.hpp", "pdac_flux.hpp"]
ARCH = os.environ.get("PDAC_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (set HIPCC)")


def build(verbose: bool = False) -> str:
    """Compile every source into LIB (atomically replaced) and return its path."""
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *FLAGS, "-I", os.path.join(ROOT, "include"), "-o", tmp,
           *[os.path.join(CSRC, s) for s in SOURCES]]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with status {res.returncode}")
    os.replace(tmp, LIB)
    return LIB


FORTRAN = os.path.join(HERE, "fortran")
DRIVER = os.path.join(FORTRAN, "pdac_drive.x")


def flang() -> str | None:
    for cand in (os.environ.get("FC"), "/opt/rocm/llvm/bin/flang", "/opt/rocm/lib/llvm/bin/flang",
                 shutil.which("flang")):
        if cand and os.path.exists(cand):
            return cand
    return None


def build_fortran(verbose: bool = False) -> str | None:
    """Compile the iso_c_binding module (fortran/pdac_gpu.f90) and the
    main_prog-shaped driver (fortran/pdac_drive.f90) against LIB with the
    image's flang; returns the driver's path, or None without flang."""
    fc = flang()
    if fc is None:
        return None
    work = os.path.join(FORTRAN, "_build")
    os.makedirs(work, exist_ok=True)
    obj = os.path.join(work, "pdac_gpu.o")
    tmp = DRIVER + ".tmp"
    cmds = [[fc, "-O2", "-c", os.path.join(FORTRAN, "pdac_gpu.f90"), "-J", work, "-o", obj],
            [fc, "-O2", "-J", work, os.path.join(FORTRAN, "pdac_drive.f90"), obj, "-L", HERE, "-lpdacgpu",
             "-Wl,-rpath,$ORIGIN/..", "-o", tmp]]
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"flang failed with status {res.returncode}")
    os.replace(tmp, DRIVER)
    return DRIVER


def stale() -> bool:
    if not os.path.exists(LIB):
        return True
    t = os.path.getmtime(LIB)
    return any(os.path.getmtime(os.path.join(CSRC, d)) > t for d in DEPS) or \
        os.path.getmtime(os.path.join(ROOT, "include", "pdacgpu.h")) > t


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_fortran(verbose=True))
