"""Builds liblbm_hip.so in-tree from lbm_amd/csrc with hipcc for gfx950.

    python -m lbm_amd.build [-v]

Every HIP translation unit is compiled with -Xclang -target-feature -Xclang -packed-fp32-ops
(no packed FP32 instructions).  The objects go to lbm_amd/_build/, the library to
lbm_amd/liblbm_hip.so.  When amdflang is present the Fortran binding module and driver
(lbm_amd/fortran) are built too: _build/lbm_driver_f32 and _build/lbm_driver_f64.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblbm_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")
FORTRAN = os.path.join(HERE, "fortran")
ARCH = "gfx950"

SOURCES = ["kernels_f32.hip", "kernels_f64.hip", "kernels_f16s.hip", "kernels_bf16s.hip", "kernels_f64t.hip", "api.cpp",
           "ensemble.cpp", "adjoint.cpp", "turbulence.hip", "turbines.cpp"]
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-I", CSRC, "-I", os.path.join(os.path.dirname(HERE), "include")]


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        out = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
        if p.returncode != 0:
            failed.append(src)
            print(out, file=sys.stderr)
        elif verbose and out.strip():
            print(out)
    if failed:
        raise RuntimeError("hipcc failed for " + ", ".join(failed))
    # calibration copy kernel (plain copy of a known size), a separate executable
    subprocess.run([HIPCC] + FLAGS + ["-o", os.path.join(BUILD, "copy_bw"), os.path.join(CSRC, "copy_bw.hip")],
                   check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    # host-only check of the per-cell adjoint against differences of the collision (no GPU use)
    subprocess.run([HIPCC] + FLAGS + ["-x", "hip", "-o", os.path.join(BUILD, "adjoint_check"),
                                      os.path.join(CSRC, "adjoint_check.cpp")],
                   check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    # host-only check of the buffer layouts (layout_check.cpp)
    subprocess.run([HIPCC] + FLAGS + ["-x", "hip", "-o", os.path.join(BUILD, "layout_check"),
                                      os.path.join(CSRC, "layout_check.cpp")],
                   check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    if os.path.exists(FLANG):
        build_fortran(verbose)
    return LIB


def build_fortran(verbose: bool = False) -> list:
    """The bind(C) module and the driver, FP32 and FP64 (-DDOUBLE_PRECISION), linked against
    liblbm_hip.so with an rpath to lbm_amd/."""
    out = []
    for prec, defs in (("f32", []), ("f64", ["-DDOUBLE_PRECISION"])):
        mdir = os.path.join(BUILD, "fmod_" + prec)
        os.makedirs(mdir, exist_ok=True)
        exe = os.path.join(BUILD, "lbm_driver_" + prec)
        cmd = [FLANG, "-cpp", "-O2"] + defs + ["-module-dir", mdir,
               os.path.join(FORTRAN, "lbm_hip_mod.F90"), os.path.join(FORTRAN, "lbm_driver.F90"),
               "-L", HERE, "-llbm_hip", "-Wl,-rpath,$ORIGIN/..", "-o", exe]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True, cwd=mdir)
        out.append(exe)
    return out


if __name__ == "__main__":
    print(build("-v" in sys.argv))
