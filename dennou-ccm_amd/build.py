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
           "sht_tune.hip", "sht_remap.hip", "sht_zonal.hip", "sht_exchange.hip",
           "sht_noalloc.hip"]


# host-only C++ (no HIP), compiled by the host compiler
HOST_SOURCES = ["sht_host.cpp", "sht_restart.cpp"]
CXX = os.environ.get("CXX", "g++")
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fvisibility=hidden", "-Wall", f"-I{os.path.join(ROOT, 'include')}"]


def _compile(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    if src.endswith(".cpp"):
        cmd = [CXX, *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", out]
    else:
        cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return out


FORTRAN = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")
FDIR = os.path.join(HERE, "fortran")
FTESTS = os.path.join(ROOT, "tests", "fortran")
FMODULES = ("wa_module", "l_module")


def build_fortran(verbose: bool = False) -> str | None:
    """Compiles the Fortran drop-in modules (wa_module, l_module; ISO_C_BINDING
    over the C ABI) into lib/<module>.o + .mod and links each drop-in test
    program lib/test_<module> against libsht_hip.so.  Skipped without amdflang."""
    if not os.path.exists(FORTRAN):
        return None
    libdir = os.path.dirname(LIB)
    steps = []
    for mod in FMODULES:
        obj = os.path.join(libdir, f"{mod}.o")
        exe = os.path.join(libdir, f"test_{mod}")
        steps += [
            [FORTRAN, "-O2", "-fPIC", "-c", os.path.join(FDIR, f"{mod}.f90"), "-module-dir", libdir, "-o", obj],
            [FORTRAN, "-O2", os.path.join(FTESTS, f"test_{mod}.f90"), obj, "-module-dir", libdir, "-I", libdir,
             "-L", libdir, "-lsht_hip", "-Wl,-rpath,$ORIGIN", "-o", exe],
        ]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True, cwd=libdir)
        if r.returncode != 0:
            raise RuntimeError(f"Fortran build failed: {' '.join(cmd)}\n{r.stderr}")
    if verbose:
        print("built the Fortran drop-in modules and their test programs in", libdir)
    return libdir


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES + HOST_SOURCES))
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
