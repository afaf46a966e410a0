#!/usr/bin/env python3
"""Build recipe for the REFERENCE implementation (geirev/LBM, Fortran) used as a checker.

TEST INFRASTRUCTURE ONLY.  Nothing in the product path (lbm_amd/, include/) may import,
link or execute anything under oracle/.  Only tests/, __graft_entry__.smoke() and the
``cpu_baseline`` leg of bench.py use it, and only as the checker or the CPU baseline.

What it does
------------
Compiles the reference's CPU path (serial or OpenMP, FP32 or FP64, D3Q19 or D3Q27) with
ROCm ``amdflang`` straight from the sources where they lie under /root/reference/src, and
writes the executables to ``oracle/_ref/bin/`` (objects in ``oracle/_ref/build/``).
Nothing from the reference is copied into the repository.

* The grid is a compile-time parameter of the reference (src/mod_dimensions.F90:21-26);
  the reference's README asks users to edit that module.  This recipe instead writes its
  own eight-line ``mod_dimensions`` module (same parameter names) into the build folder
  and compiles it in place of the reference's copy.
* Flags follow src/makefile:82-113 (``-DDOUBLE_PRECISION -fdefault-real-8`` for DP,
  ``-fopenmp -DOPEN_MP`` for MP, ``-DD3Q19``) with amdflang's spelling.
* The cylinder of ``experiment cylinder`` is hard-coded (radius 7 at i=50, j=nyg/2,
  src/m_cylinder.F90:8).  With ``cylinder=(ipos, jpos, radius)`` the recipe writes its own
  ``m_cylinder`` module (same interface, same inside test ``(i-ipos)**2+(j-jpos)**2 <
  radius**2`` over interior i, j and every k) with those parameters and compiles it in place
  of the reference's copy, as it does for ``mod_dimensions``.
* The MPI modules (src/m_mpi_*.F90) and NetCDF output are left out (no -DMPI/-DNETCDF).
* FFTW3 is absent from the image.  Only the inflow-turbulence generator calls it
  (src/m_pseudo2D.F90:78,150), so the link leaves those three symbols unresolved
  (``--unresolved-symbols=ignore-all``); runs driven by tests/golden never reach them
  because turbulence fields are injected through restart files with nrturb > nt1.

Usage: python oracle/build_ref.py --q 19 --prec sp --dims 16 8 6 [--omp]
"""
import argparse
import os
import re
import subprocess
import sys

REF_SRC = "/root/reference/src"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")
FC = "/opt/rocm/bin/amdflang"


def variant_name(q, prec, dims, omp, cylinder=None):
    nx, ny, nz = dims
    cyl = f"_cyl{cylinder[0]}x{cylinder[1]}r{cylinder[2]}" if cylinder else ""
    return f"q{q}_{prec}_{nx}x{ny}x{nz}{cyl}{'_omp' if omp else ''}"


def exe_path(q, prec, dims, omp=False, cylinder=None):
    return os.path.join(OUT, "bin", variant_name(q, prec, dims, omp, cylinder))


def _sources(cylinder=None):
    files = sorted(f for f in os.listdir(REF_SRC) if f.endswith(".F90"))
    skip = {"mod_dimensions.F90"} | ({"m_cylinder.F90"} if cylinder else set())
    return [f for f in files if f not in skip and not f.startswith("m_mpi_")]


def _topo_order(files, extra_modules):
    """Order sources so that every module is compiled before its users."""
    provides, needs = {}, {}
    mod_re = re.compile(r"^\s*module\s+(\w+)\s*$", re.I)
    use_re = re.compile(r"^\s*use\s*(?:,\s*intrinsic\s*::)?\s*(\w+)", re.I)
    for f in files:
        mods, uses = set(), set()
        with open(os.path.join(REF_SRC, f), errors="replace") as fh:
            for line in fh:
                m = mod_re.match(line)
                if m and m.group(1).lower() != "procedure":
                    mods.add(m.group(1).lower())
                u = use_re.match(line)
                if u:
                    uses.add(u.group(1).lower())
        needs[f] = uses
        for m in mods:
            provides[m] = f
    order, done, visiting = [], set(), set()

    def visit(f):
        if f in done:
            return
        if f in visiting:
            raise RuntimeError(f"module cycle at {f}")
        visiting.add(f)
        for u in sorted(needs[f]):
            if u in provides and provides[u] != f:
                visit(provides[u])
        visiting.discard(f)
        done.add(f)
        order.append(f)

    for f in files:
        visit(f)
    return order


def build(q, prec, dims, omp=False, verbose=False, cylinder=None):
    name = variant_name(q, prec, dims, omp, cylinder)
    exe = exe_path(q, prec, dims, omp, cylinder)
    if os.path.exists(exe):
        return exe
    if not os.path.isdir(REF_SRC):
        raise FileNotFoundError("reference sources not present; cannot build " + name)
    bdir = os.path.join(OUT, "build", name)
    os.makedirs(bdir, exist_ok=True)
    os.makedirs(os.path.dirname(exe), exist_ok=True)
    nx, ny, nz = dims
    dimsrc = os.path.join(bdir, "mod_dimensions.F90")
    with open(dimsrc, "w") as fh:
        fh.write(
            "module mod_dimensions\n"
            f"   integer, parameter :: nx = {nx}\n"
            f"   integer, parameter :: ny = {ny}\n"
            "   integer, parameter :: ntiles = 1\n"
            "   integer, parameter :: nyg = ntiles*ny\n"
            f"   integer, parameter :: nz = {nz}\n"
            "   integer, parameter :: ntracer = 0\n"
            "end module\n")
    own = [dimsrc]
    if cylinder:
        ipos, jpos, radius = (int(x) for x in cylinder)
        cylsrc = os.path.join(bdir, "m_cylinder.F90")
        with open(cylsrc, "w") as fh:
            fh.write(
                "module m_cylinder\n"
                "contains\n"
                "subroutine cylinder(blanking)\n"
                "   use mod_dimensions, only : nx, nyg, nz\n"
                "   implicit none\n"
                "   logical, intent(inout) :: blanking(0:nx+1,0:nyg+1,0:nz+1)\n"
                "   integer :: ii, jj\n"
                f"   integer, parameter :: r = {radius}, ic = {ipos}, jc = {jpos}\n"
                "   blanking = .false.\n"
                "   do jj = 1, nyg\n"
                "      do ii = 1, nx\n"
                "         blanking(ii,jj,:) = (ii-ic)**2 + (jj-jc)**2 < r**2\n"
                "      end do\n"
                "   end do\n"
                "end subroutine\n"
                "end module\n")
        own.append(cylsrc)
    flags = ["-cpp", "-O2"]
    if q == 19:
        flags.append("-DD3Q19")
    if prec == "dp":
        flags += ["-DDOUBLE_PRECISION", "-fdefault-real-8"]
    if omp:
        flags += ["-fopenmp", "-DOPEN_MP"]
    objs = []
    # the project-written modules first: every reference module that uses them comes later
    srcs = own + [os.path.join(REF_SRC, f) for f in _topo_order(_sources(cylinder), {"mod_dimensions"})]
    for src in srcs:
        obj = os.path.join(bdir, os.path.basename(src)[:-4] + ".o")
        cmd = [FC] + flags + ["-module-dir", bdir, "-I", bdir, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd))
        subprocess.run(cmd, check=True, cwd=bdir, stdout=subprocess.DEVNULL if not verbose else None)
        objs.append(obj)
    link = [FC] + (["-fopenmp"] if omp else []) + objs + ["-o", exe, "-Wl,--unresolved-symbols=ignore-all"]
    subprocess.run(link, check=True, cwd=bdir)
    return exe


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--q", type=int, choices=(19, 27), required=True)
    ap.add_argument("--prec", choices=("sp", "dp"), default="sp")
    ap.add_argument("--dims", type=int, nargs=3, required=True)
    ap.add_argument("--omp", action="store_true")
    ap.add_argument("--cylinder", type=int, nargs=3, default=None, metavar=("IPOS", "JPOS", "RADIUS"))
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.q, a.prec, tuple(a.dims), a.omp, a.verbose, tuple(a.cylinder) if a.cylinder else None))


if __name__ == "__main__":
    sys.exit(main())
