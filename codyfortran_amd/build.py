"""Build the gfx950 HIP library ``codyfortran_amd/libcody_hip.so`` in-tree with hipcc.

Usage: ``python -m codyfortran_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcody_hip.so")

SOURCES = ["host_setup.cpp", "state_files.cpp", "batch_plan.cpp", "mctdhx_kernels.hip", "ci_kernels.hip", "meanfield_gemm.hip",
           "meanfield_lowprec.hip",
           "cody_api.hip", "cody_propagate.hip", "cody_adjoint.hip", "noalloc.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# The packed-FP32 feature is switched off for every device pass (hipcc forms it by itself from
# plain float code; some of its forms are not trusted on this hardware).
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-I", os.path.join(ROOT, "include"),
]


def _compile(src: str) -> str:
    os.makedirs(OBJ, exist_ok=True)
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC, *FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return out


FORTRAN_DIR = os.path.join(HERE, "fortran")
FORTRAN_SOURCES = ["M_Utils_Types.f90", "M_CodyHip.f90", "M_Method.f90", "M_Propagator.f90",
                   "M_Grid.f90", "M_Method_Mb_OrbBased.f90", "M_Method_Mb_OrbBased_MctdhxHip.f90",
                   "S_Method_Mb_OrbBased_MctdhxHip.f90", "M_IntegratorList.f90", "M_IntegratorList_Hip.f90",
                   "S_IntegratorList_Hip.f90", "S_IntegratorList.f90"]
# test programs: (source, executable) -- each links the library and the objects above
FORTRAN_DRIVERS = [
    (os.path.join(ROOT, "tests", "fortran", "T_MctdhxHip_Rhs.f90"), "t_mctdhx_hip"),
    (os.path.join(ROOT, "tests", "fortran", "T_StateFiles.f90"), "t_state_files"),
    (os.path.join(ROOT, "tests", "fortran", "T_MctdhxHip_Split.f90"), "t_mctdhx_split"),
]
BIN = os.path.join(HERE, "bin")
FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")


def _fc(path: str, mod: str) -> str:
    out = os.path.join(mod, os.path.splitext(os.path.basename(path))[0] + ".o")
    cmd = [FC, "-O2", "-fPIC", "-module-dir", mod, "-I", mod, "-c", path, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"amdflang failed for {path}:\n{r.stdout}\n{r.stderr}")
    return out


def build_fortran(verbose: bool = False) -> list[str]:
    """Fortran host API (mirror modules + GPU leaf) and the test programs."""
    mod = os.path.join(OBJ, "fortran")
    os.makedirs(mod, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    objs = [_fc(os.path.join(FORTRAN_DIR, src), mod) for src in FORTRAN_SOURCES]
    exes = []
    for src, name in FORTRAN_DRIVERS:
        exe = os.path.join(BIN, name)
        cmd = [FC, *objs, _fc(src, mod), "-o", exe, "-L", HERE, "-lcody_hip", "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"Fortran link failed for {name}:\n{r.stdout}\n{r.stderr}")
        if verbose:
            print(f"built {exe}")
        exes.append(exe)
    return exes


def build_library(verbose: bool = False) -> str:
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build_library(verbose=True)
    build_fortran(verbose=True)
    sys.exit(0)
