"""Build the MI355X RK3-step library (and the Fortran ISO_C_BINDING shim) in tree.

    python pandora-2.0_amd/build_lib.py

Outputs (git-ignored, shipped with the source tree):
    pandora-2.0_amd/lib/libpandora_amd.so     HIP kernels + C-ABI (include/pandora_amd.h)
    pandora-2.0_amd/lib/pdr_rk3_amd.o/.mod    Fortran shim module (if amdflang is present)
    pandora-2.0_amd/lib/pdr_fortran_driver    Fortran driver program using the shim
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
FSRC = os.path.join(PKG, "fortran")
LIB = os.path.join(PKG, "lib")
SO = os.path.join(LIB, "libpandora_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("PDR_FC", "/opt/rocm/bin/amdflang")
ARCH = "gfx950"

SOURCES = ["pdr_api.hip", "pdr_kernels.hip", "pdr_kernels2.hip", "pdr_host.hip", "pdr_stats.hip", "pdr_pstats.hip", "pdr_sort.hip", "pdr_ens.hip"]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_hip(verbose: bool = True) -> str:
    """Compile every HIP source (in parallel, one hipcc per source) and link the library."""
    from concurrent.futures import ThreadPoolExecutor

    os.makedirs(LIB, exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        obj = os.path.join(LIB, os.path.splitext(src)[0] + ".o")
        cmds.append([HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-c",
                     # no packed-FP32 instructions: some of their forms have given wrong results on gfx950
                     "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                     "-Wall", "-Wno-unused-parameter", "-Wno-unused-command-line-argument",
                     "-I", os.path.join(ROOT, "include"), "-I", CSRC,
                     os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "0") or 0) or (os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(_run, cmd) for cmd in cmds]:
            f.result()
    tmp = SO + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, SO)
    return SO


def build_fortran() -> bool:
    """Compile the ISO_C_BINDING shim and its driver with amdflang, if present."""
    if not os.path.exists(FC) or not os.path.isdir(FSRC):
        print(f"[build] Fortran compiler {FC} not found; shim not built", flush=True)
        return False
    os.makedirs(LIB, exist_ok=True)
    shim = os.path.join(FSRC, "pdr_rk3_amd.f90")
    drv = os.path.join(FSRC, "pdr_fortran_driver.f90")
    _run([FC, "-O2", "-fPIC", "-c", shim, "-o", os.path.join(LIB, "pdr_rk3_amd.o"),
          "-module-dir", LIB])
    if os.path.exists(drv):
        _run([FC, "-O2", drv, os.path.join(LIB, "pdr_rk3_amd.o"), "-I", LIB,
              "-L", LIB, "-lpandora_amd", "-Wl,-rpath,$ORIGIN",
              "-o", os.path.join(LIB, "pdr_fortran_driver")])
    return True


def build() -> None:
    build_hip()
    build_fortran()


if __name__ == "__main__":
    build_hip()
    if "--no-fortran" not in sys.argv:
        build_fortran()
