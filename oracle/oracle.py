"""ctypes front end of the C restatement (oracle/lbm_oracle.c) and runner of the reference
build (oracle/build_ref.py).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product path.

``OracleSim`` mirrors the reference driver's restart path and time loop:
  readrestart -> fA; macrovars(fA); fB = fA; tau = 3*kinevisc + 0.5   (src/main.F90:255-262)
  do it = nt0+1, nt1: the step sequence of src/main.F90:290-329       (or_run)
"""
from __future__ import annotations

import ctypes as C
import os
import shutil
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


class _Params(C.Structure):
    pass


def _params_type(real):
    class P(C.Structure):
        _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("nl", C.c_int),
                    ("ibgk", C.c_int), ("ihrr", C.c_int), ("ivreman", C.c_int),
                    ("ibnd", C.c_int), ("jbnd", C.c_int), ("kbnd", C.c_int),
                    ("iturb_pos", C.c_int),
                    ("kinevisc", real), ("smagorinsky", real), ("rho0", real)]
    return P


_LIBS = {}

# NREL 5-MW blade sections (src/mod_nrel5mw.F90:24-48): radial position, segment length and chord
# [m], twist [deg], airfoil index (1-based into NREL5MW_FOILS); rotor and hub radius [m].
NREL5MW_RELM = [2.86670, 5.60000, 8.33330, 11.75000, 15.85000, 19.95000, 24.05000, 28.15000, 32.25000,
                36.35000, 40.45000, 44.55000, 48.65000, 52.75000, 56.16670, 58.90000, 61.63330]
NREL5MW_DC = [2.73330, 2.73330, 2.73330] + [4.10000] * 11 + [2.73330, 2.73330, 2.73330]
NREL5MW_CHORD = [3.54200, 3.85400, 4.16700, 4.55700, 4.65200, 4.45800, 4.24900, 4.00700, 3.74800, 3.50200,
                 3.25600, 3.01000, 2.76400, 2.51800, 2.31300, 2.08600, 1.41900]
NREL5MW_TWIST = [13.3080, 13.3080, 13.3080, 13.3080, 11.4800, 10.1620, 9.0110, 7.7950, 6.5440, 5.3610, 4.1880,
                 3.1250, 2.3190, 1.5260, 0.8630, 0.3700, 0.1060]
NREL5MW_NFOIL = [1, 1, 2, 3, 4, 4, 5, 6, 6, 7, 7, 8, 8, 8, 8, 8, 8]
NREL5MW_FOILS = ["Cylinder1.dat", "Cylinder2.dat", "DU40_A17.dat", "DU35_A17.dat", "DU30_A17.dat", "DU25_A17.dat",
                 "DU21_A17.dat", "NACA64_A17.dat"]   # src/m_nrelreadfoil.F90:13-20


def read_foil_text(text, prec="f32"):
    """One airfoil table as src/m_nrelreadfoil.F90:33-47 reads it: 50 lines skipped, the row count,
    2 lines skipped, then rows alpha [deg], cl, cd, cm.  Returns an (n, 3) array of deg, cl, cd."""
    R = np.float32 if prec == "f32" else np.float64
    lines = text.splitlines()
    n = int(lines[50].split()[0])
    rows = [[R(x) for x in lines[53 + r].replace(",", " ").split()[:3]] for r in range(n)]
    return np.array(rows, R)


def turbines_from_infile(pos, rpm, pitch, length, vel, foils, prec="f32", theta=None, nblades=3):
    """turbine_initialize (src/m_turbine_initialize.F90:29-70) on infile values: pos = [(ipos, jpos,
    kpos, yaw_deg, tilt_deg)], rpm, pitch [deg], C_l = length [m], C_u = vel [m/s]; foils = list of
    (n, 3) tables in NREL5MW_FOILS order.  Every conversion rounds in the REAL kind.  Returns the dict
    OracleSim.run(turbines=...) takes; theta (rad) defaults to zeros."""
    R = np.float32 if prec == "f32" else np.float64
    pi = R(3.14159265358979323846)
    pi2 = R(2.0) * pi
    length, vel = R(length), R(vel)
    p2l_time = length / vel
    omega = pi2 * R(rpm) / R(60.0)
    n = len(pos)
    return dict(
        xhub=np.array([R(p[0]) for p in pos], R), yhub=np.array([R(p[1]) for p in pos], R),
        zhub=np.array([R(p[2]) for p in pos], R),
        yaw=np.array([(R(p[3]) / R(360.0)) * pi2 for p in pos], R),
        tilt=np.array([(R(p[4]) / R(360.0)) * pi2 for p in pos], R),
        omegand=np.full(n, omega * p2l_time, R), pitch=np.full(n, R(pitch), R),
        theta=np.zeros(n, R) if theta is None else np.asarray(theta, R).copy(),
        relm=np.array([R(x) / length for x in NREL5MW_RELM], R), dc=np.array([R(x) / length for x in NREL5MW_DC], R),
        chord=np.array([R(x) / length for x in NREL5MW_CHORD], R), twist=np.array(NREL5MW_TWIST, R),
        nfoil=np.array(NREL5MW_NFOIL, np.int32), foils=[np.asarray(f, R) for f in foils], nblades=nblades)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _lib(prec):
    if prec not in _LIBS:
        path = os.path.join(HERE, f"liblbm_oracle_{prec}.so")
        if not os.path.exists(path):
            build()
        _LIBS[prec] = C.CDLL(path)
    return _LIBS[prec]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _turbines_struct(real, T, keep):
    """ctypes image of or_turbines for the dict T (arrays kept alive in ``keep``)."""
    class S(C.Structure):
        _fields_ = [(n, C.c_int) for n in ("nturb", "nblades", "nchords", "nfoils", "nang_max")] + \
                   [(n, C.c_void_p) for n in ("xhub", "yhub", "zhub", "yaw", "tilt", "omegand", "pitch", "theta",
                                              "relm", "dc", "chord", "twist", "nfoil", "nang", "foil")]
    dt = np.float32 if real is C.c_float else np.float64
    nmax = max(len(f) for f in T["foils"])
    foil = np.zeros((len(T["foils"]), nmax, 3), dt)
    for k, f in enumerate(T["foils"]):
        foil[k, :len(f)] = f
    arr = {k: np.ascontiguousarray(T[k], dt) for k in ("xhub", "yhub", "zhub", "yaw", "tilt", "omegand", "pitch",
                                                       "relm", "dc", "chord", "twist")}
    arr["theta"] = T["theta"]            # advanced in place
    assert arr["theta"].dtype == dt and arr["theta"].flags.c_contiguous
    arr["nfoil"] = np.ascontiguousarray(T["nfoil"], np.int32)
    arr["nang"] = np.array([len(f) for f in T["foils"]], np.int32)
    arr["foil"] = foil
    keep.extend(arr.values())
    s = S(len(T["xhub"]), int(T.get("nblades", 3)), len(T["relm"]), len(T["foils"]), nmax,
          *[arr[k].ctypes.data for k in ("xhub", "yhub", "zhub", "yaw", "tilt", "omegand", "pitch", "theta",
                                          "relm", "dc", "chord", "twist", "nfoil", "nang", "foil")])
    keep.append(s)
    return s


class OracleSim:
    """Reference-semantics simulation on the CPU (FP32 or FP64)."""

    def __init__(self, q, dims, prec="f32", ibgk=2, ihrr=0, ivreman=0, smagorinsky=0.0,
                 kinevisc=1.0 / 30.0, ibnd=0, jbnd=0, kbnd=0, rho0=1.0, iturb_pos=10):
        self.prec = prec
        self.dt = np.float32 if prec == "f32" else np.float64
        real = C.c_float if prec == "f32" else C.c_double
        self.lib = _lib(prec)
        P = _params_type(real)
        nx, ny, nz = dims
        self.q, self.dims = q, (nx, ny, nz)
        self.p = P(nx, ny, nz, q, ibgk, ihrr, ivreman, ibnd, jbnd, kbnd, iturb_pos,
                   kinevisc, smagorinsky, rho0)
        self.kinevisc = self.dt(kinevisc)
        shp = (nz + 2, ny + 2, nx + 2)
        self.fA = np.zeros(shp + (q,), self.dt)
        self.fB = np.zeros(shp + (q,), self.dt)
        self.rho, self.u, self.v, self.w = (np.zeros(shp, self.dt) for _ in range(4))
        self.tau = np.zeros(shp, self.dt)
        self.cur = 0   # 0: f1 is fA
        self.lib.or_run.restype = C.c_int

    # -- state ------------------------------------------------------------------------
    @property
    def f1(self):
        return self.fA if self.cur == 0 else self.fB

    @property
    def f2(self):
        return self.fB if self.cur == 0 else self.fA

    def restart(self, f):
        """readrestart + macrovars + fB=fA + tau (src/main.F90:255-262)."""
        self.fA[...] = f
        self.fB[...] = f
        self.cur = 0
        self.lib.or_macrovars(C.byref(self.p), _ptr(self.fA), _ptr(self.rho), _ptr(self.u),
                              _ptr(self.v), _ptr(self.w))
        self.tau[...] = self.dt(3.0) * self.kinevisc + self.dt(0.5)

    def init_equilibrium(self, rho, u, v, w, it0=0, uini=0.0, udir=0.0, uvel_shear=None):
        """nt0 = 0 path (src/main.F90:243-247): fequil3(fB), fillghosts(fB), fA = fB,
        boundarycond(fB,fA), boundarycond(fA,fB); the first postcoll uses these moments."""
        dt = self.dt
        for dst, a in ((self.rho, rho), (self.u, u), (self.v, v), (self.w, w)):
            dst[...] = np.asarray(a, dt)
        self.fB[...] = 0
        self.lib.or_fequil3(C.byref(self.p), _ptr(self.fB), _ptr(self.rho), _ptr(self.u), _ptr(self.v), _ptr(self.w))
        self.lib.or_fillghosts(C.byref(self.p), _ptr(self.fB))
        self.fA[...] = self.fB
        nz = self.dims[2]
        shear = np.ones(nz, dt) if uvel_shear is None else np.asarray(uvel_shear, dt)
        uvel = np.ascontiguousarray(dt(uini) * shear, dt)
        real = C.c_float if self.prec == "f32" else C.c_double
        self.lib.or_boundarycond(C.byref(self.p), _ptr(self.fB), _ptr(self.fA), _ptr(uvel), real(udir))
        self.lib.or_boundarycond(C.byref(self.p), _ptr(self.fA), _ptr(self.fB), _ptr(uvel), real(udir))
        self.cur = 0
        self.tau[...] = self.dt(3.0) * self.kinevisc + self.dt(0.5)

    def macros(self):
        return self.rho.copy(), self.u.copy(), self.v.copy(), self.w.copy()

    # -- time loop --------------------------------------------------------------------
    def run(self, it0, it1, uini=0.0, udir=0.0, uvel_shear=None, uvel_time=None, udir_time=None,
            blank=None, turb=None, forcing=None, ablvisc=None, turbines=None, buoyancy=None):
        """Steps it0+1..it1.  turbines: the dict of turbines_from_infile; its theta advances.
        buoyancy: (pottemp, scaling, theta0) with pottemp the it1 - it0 potential-temperature fields
        (nz+2, ny+2, nx+2) of the steps in order (src/main.F90:293)."""
        nx, ny, nz = self.dims
        dt = self.dt
        shear = np.ones(nz, dt) if uvel_shear is None else np.asarray(uvel_shear, dt)
        ut = np.full(it1 + 1, uini, dt) if uvel_time is None else np.asarray(uvel_time, dt)
        dtm = np.full(it1 + 1, udir, dt) if udir_time is None else np.asarray(udir_time, dt)
        abl = np.zeros(nz, dt) if ablvisc is None else np.asarray(ablvisc, dt)
        uu = vv = ww = None
        nrturb, ampl = 1, 0.0
        if turb is not None:
            uu, vv, ww, nrturb, ampl = turb
            uu, vv, ww = (np.ascontiguousarray(a, dt) for a in (uu, vv, ww))
        F = None if forcing is None else np.ascontiguousarray(forcing, dt)
        bl = None if blank is None else np.ascontiguousarray(blank, np.uint8)
        real = C.c_float if self.prec == "f32" else C.c_double
        keep = []
        T = None if turbines is None else C.byref(_turbines_struct(real, turbines, keep))
        pot, bscale, btheta0 = None, 0.0, 1.0
        if buoyancy is not None:
            pot, bscale, btheta0 = buoyancy
            pot = np.ascontiguousarray(pot, dt)
            if pot.shape != (it1 - it0, nz + 2, ny + 2, nx + 2):
                raise ValueError(f"buoyancy fields: {pot.shape}, expected {(it1 - it0, nz + 2, ny + 2, nx + 2)}")
        sw = self.lib.or_run(C.byref(self.p), _ptr(self.f1), _ptr(self.f2), _ptr(self.tau),
                             _ptr(self.rho), _ptr(self.u), _ptr(self.v), _ptr(self.w), _ptr(abl),
                             _ptr(bl), _ptr(shear), _ptr(ut), _ptr(dtm), _ptr(uu), _ptr(vv), _ptr(ww),
                             C.c_int(nrturb), real(ampl), _ptr(F), C.c_int(F is not None), C.c_int(it0),
                             C.c_int(it1), T, _ptr(pot), real(bscale), real(btheta0))
        self.cur ^= sw
        return self.f1


def inflow_series(files, nz, nt0, nt1, uini, udir, prec="f32"):
    """src/m_inflow.F90:27-86 on the texts of uvel_shear.dat / uvel_time.dat (``files``: name ->
    text, either may be absent) with p2l%vel = p2l%time = 1 (C_l = C_u = 1 in infile_text).
    Returns (uvel_shear[nz], uvel_time[0..nt1], udir_time[0..nt1]) indexed by absolute step, the
    form OracleSim.run takes; every operation rounds in the build's REAL kind."""
    R = np.float32 if prec == "f32" else np.float64
    vals = lambda txt: [[R(x) for x in line.replace(",", " ").split()] for line in txt.splitlines() if line.strip()]
    uvel_h = np.ones(nz, R)
    if "uvel_shear.dat" in files:
        for k, row in enumerate(vals(files["uvel_shear.dat"])[:nz]):
            uvel_h[k] = row[2]                        # read(10,*) kk, z(k), uvel_h(k)
    shear = np.array([uvel_h[k] / uvel_h[nz - 1] for k in range(nz)], R)
    ut = np.full(nt1 + 1, R(uini), R)
    ud = np.full(nt1 + 1, R(udir), R)
    if "uvel_time.dat" in files:
        rows = vals(files["uvel_time.dat"])
        tdata = [r[0] for r in rows]
        utd = [r[1] / R(1.0) for r in rows]            # uvel_tdata(k) / p2l%vel
        udd = [r[2] for r in rows]
        last = len(rows) - 1
        for i in range(nt0, nt1 + 1):
            t = R(i - 1) * R(1.0)                       # real(i-1)*p2l%time
            if t <= tdata[0]:
                ut[i], ud[i] = utd[0], udd[0]
            elif t >= tdata[last]:
                ut[i], ud[i] = utd[last], udd[last]
            else:
                k = next(k for k in range(last) if tdata[k] <= t <= tdata[k + 1])
                tmp = (t - tdata[k]) / (tdata[k + 1] - tdata[k])
                ut[i] = utd[k] + tmp * (utd[k + 1] - utd[k])
                ud[i] = udd[k] + tmp * (udd[k + 1] - udd[k])
    return shear, ut, ud


# ------------------------------------------------------------------------------------------
# Reference build runner
# ------------------------------------------------------------------------------------------
def infile_text(*, ibgk=2, ihrr=0, ivreman=0, smagorinsky=0.0, nt0=1, nt1=2, ibnd=0, jbnd=0, kbnd=0,
                uini=0.0, udir=0.0, turb=False, ampl=0.0, nrturb=100, visc=1.0 / 30.0,
                experiment="none", ltesting=True, ldump=True, ltiming=False, iout=None,
                length=1.0, vel=1.0, turbines=None, iablvisc=0, ablheight=400.0, istable=1):
    """An infile.in in the line order src/m_readinfile.F90:84-160 reads.  With the default
    C_l = C_u = 1 the viscosity line is the non-dimensional kinevisc (src/m_readinfile.F90:182-186)
    and uini the lattice speed; ``length``/``vel`` set C_l [m] and C_u [m/s] (then the caller
    converts: kinevisc = visc / (C_l * C_u), lattice uini = uini / C_u).  ``turbines``: dict with
    pitch [deg], rpm, tsr, tiploss and pos = [(ipos, jpos, kpos, yaw, tilt), ...]
    (src/m_readinfile.F90:143-161).  iablvisc, ablheight, istable: the atmospheric boundary layer
    section (src/m_readinfile.F90:101-103)."""
    iout = iout if iout is not None else 10 ** 7
    L = lambda b: "T" if b else "F"
    smagorinsky, uini, udir, ampl, visc = (float(x) for x in (smagorinsky, uini, udir, ampl, visc))
    rows = [
        "# Development variables",
        f" {L(ltiming)}", f" {L(ltesting)}", f" {L(ldump)}", " 64 2 1", " F",
        "# Experiment configuration",
        f" {experiment}", f" {ibgk}", f" {ihrr}", f" {ivreman} {smagorinsky!r}",
        "# Atmospheric boundary layer",
        f" {int(iablvisc)}", f" {float(ablheight)!r}", f" {int(istable)}",
        "# Time stepping and outputs",
        f" {nt0}", f" {nt1}", f" {iout}", f" {10 ** 7}", f" 0 {10 ** 7} {10 ** 7}", " 2",
        "# Boundary conditions",
        f" {ibnd}", f" {jbnd}", f" {kbnd}",
        "# Inflow variables",
        f" {uini!r} {udir!r}", f" {L(turb)} {ampl!r} {nrturb}",
        "# Physical variables",
        f" {visc!r}", " 1.225", f" {float(length)!r}", f" {float(vel)!r}",
        "# Averaging variables",
        " F F", " 20000", " 40000",
        "# Turbine-definitions",
    ]
    if turbines:
        t = turbines
        rows += [f" {len(t['pos'])}", f" {float(t.get('pitch', 0.0))!r}", f" {float(t['rpm'])!r}",
                 f" {float(t.get('tsr', 0.0))!r}", f" {int(t.get('tiploss', 0))}"]
        for n, (ip, jp, kp, yw, tl) in enumerate(t["pos"]):
            rows += [f"# turbine {n + 1}", f" {ip} {jp} {kp} {float(yw)!r} {float(tl)!r}"]
    else:
        rows += [" 0"]
    return "\n".join(rows) + "\n"


def run_reference(exe, f_in, *, nt0, nt1, turb_fields=None, workdir=None, threads=1, timeout=3600,
                  read_output=True, files=None, theta=None, pottemp=None, **opts):
    """Run a reference executable (built by oracle/build_ref.py) from restart state ``f_in``
    at step nt0 to nt1.  Returns (f_out, stdout).  f_out is the restart record written at
    nt1 (post-streaming f1 with ghosts, src/main.F90:398).  ``files``: name -> text of input
    files written into the working directory first (uvel_shear.dat, uvel_time.dat, Airfoils/*.dat).
    ``theta``: the turbines' azimuths at nt0 (runs with turbines need that restart record); with it
    the result is (f_out, theta_out, stdout).  ``pottemp``: the potential temperature at nt0 (runs with
    iablvisc = 2 need that restart record, src/m_readrestart.F90:119-136); with it the result is
    (f_out, pottemp_out, stdout), pottemp_out the record the reference writes at nt1."""
    import sys
    sys.path.insert(0, os.path.dirname(HERE))
    from lbm_amd import fortran_io as fio
    own = workdir is None
    wd = tempfile.mkdtemp(prefix="lbmref_") if own else workdir
    try:
        os.makedirs(os.path.join(wd, "restart"), exist_ok=True)
        fio.write_restart(os.path.join(wd, "restart", fio.restart_name(nt0)), f_in)
        if theta is not None:
            fio.write_theta(os.path.join(wd, "restart", fio.restart_name(nt0, prefix="theta")), theta,
                            np.asarray(f_in).dtype)
        if pottemp is not None:
            fio.write_pottemp(os.path.join(wd, "restart", fio.restart_name(nt0, prefix="pottemp")), pottemp)
        if turb_fields is not None:
            uu, vv, ww, rr = turb_fields
            fio.write_turbulence(os.path.join(wd, "restart", fio.restart_name(nt0, prefix="turbulence")),
                                 uu, vv, ww, rr)
        for fname, text in (files or {}).items():
            os.makedirs(os.path.dirname(os.path.join(wd, fname)), exist_ok=True)
            with open(os.path.join(wd, fname), "w") as fh:
                fh.write(text)
        with open(os.path.join(wd, "infile.in"), "w") as fh:
            fh.write(infile_text(nt0=nt0, nt1=nt1, turb=turb_fields is not None, **opts))
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        p = subprocess.run(["bash", "-c", f"ulimit -s unlimited; {exe}"], cwd=wd, env=env,
                           capture_output=True, text=True, timeout=timeout)
        if p.returncode != 0 or "infile.in problem" in p.stdout:
            raise RuntimeError(f"reference run failed ({p.returncode}):\n{p.stdout[-3000:]}\n{p.stderr[-3000:]}")
        if not read_output:
            return None, p.stdout
        f_out = fio.read_restart(os.path.join(wd, "restart", fio.restart_name(nt1)))
        if theta is not None:
            th = fio.read_theta(os.path.join(wd, "restart", fio.restart_name(nt1, prefix="theta")), f_out.dtype)
            return f_out, th, p.stdout
        if pottemp is not None:
            pt = fio.read_pottemp(os.path.join(wd, "restart", fio.restart_name(nt1, prefix="pottemp")),
                                  np.asarray(pottemp).shape, f_out.dtype)
            return f_out, pt, p.stdout
        return f_out, p.stdout
    finally:
        if own:
            shutil.rmtree(wd, ignore_errors=True)
