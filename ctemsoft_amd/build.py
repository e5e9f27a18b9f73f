"""Build the HIP library ctemsoft_amd/libctem_ebsd.so for gfx950 (in-tree).

    python -m ctemsoft_amd.build

Every HIP source is compiled with packed-FP32 instructions disabled
(-target-feature -packed-fp32-ops) and linked into one shared library whose exported
symbols are exactly the C-ABI of include/ctem_ebsd.h.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libctem_ebsd.so")
SOURCES = ["ctem_kernels.hip", "ctem_mixed.hip", "ctem_split.hip", "ctem_adjoint.hip", "ctem_api.hip", "ctem_detector.hip",
           "ctem_ecp.hip", "ctem_mc.hip",
           "ctem_setup.hip"]
ARCH = os.environ.get("CTEM_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def flags() -> list[str]:
    return ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-fvisibility=hidden", f"-I{os.path.join(REPO, 'include')}"]


def build_fortran(verbose: bool = True) -> str | None:
    """ISO_C_BINDING interface module and the Fortran boundary check program, linked to the
    HIP library (amdflang from the ROCm image)."""
    fc = "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        return None
    fdir = os.path.join(HERE, "fortran")
    out = os.path.join(fdir, "_build")
    os.makedirs(out, exist_ok=True)
    exe = os.path.join(HERE, "ctem_lgh_check")
    cmd = [fc, "-O2", f"-J{out}", os.path.join(fdir, "ctem_ebsd_iface.f90"),
           os.path.join(fdir, "ctem_lgh_check.f90"), f"-L{HERE}", "-lctem_ebsd",
           "-Wl,-rpath,$ORIGIN", "-o", exe]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=out)
    return exe


def build(verbose: bool = True) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmd = [hipcc()] + flags() + ["-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
