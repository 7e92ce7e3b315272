"""Builds libsht_hip.so (gfx950) in-tree: dennou-ccm_amd/lib/libsht_hip.so.

hipcc compiles each HIP translation unit for gfx950 with packed-FP32 ops
disabled, then links one shared library whose exports are exactly the
extern "C" entry points of include/sht.h.
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
LIB = os.path.join(HERE, "lib", "libsht_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", f"-I{os.path.join(ROOT, 'include')}",
]
SOURCES = ["sht_capi.hip", "sht_tables.hip", "sht_fft.hip", "sht_legendre.hip",
           "sht_legendre_mixed.hip", "sht_pack.hip", "sht_batch.hip", "sht_server.hip",
           "sht_tune.hip", "sht_remap.hip"]


def _compile(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return out


FORTRAN = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")
FDIR = os.path.join(HERE, "fortran")
FTEST = os.path.join(ROOT, "tests", "fortran", "test_wa_module.f90")


def build_fortran(verbose: bool = False) -> str | None:
    """Compiles the Fortran wa_module (ISO_C_BINDING over the C ABI) into
    lib/wa_module.o + lib/wa_module.mod and links the drop-in test program
    lib/test_wa_module against libsht_hip.so.  Skipped without amdflang."""
    if not os.path.exists(FORTRAN):
        return None
    libdir = os.path.dirname(LIB)
    obj = os.path.join(libdir, "wa_module.o")
    exe = os.path.join(libdir, "test_wa_module")
    steps = [
        [FORTRAN, "-O2", "-fPIC", "-c", os.path.join(FDIR, "wa_module.f90"), "-module-dir", libdir, "-o", obj],
        [FORTRAN, "-O2", FTEST, obj, "-module-dir", libdir, "-I", libdir, "-L", libdir, "-lsht_hip",
         "-Wl,-rpath,$ORIGIN", "-o", exe],
    ]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True, cwd=libdir)
        if r.returncode != 0:
            raise RuntimeError(f"Fortran build failed: {' '.join(cmd)}\n{r.stderr}")
    if verbose:
        print("built", exe)
    return exe


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
