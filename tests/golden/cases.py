"""Golden-vector cases: the reference's own build (oracle/build_ref.py) run on seeded inputs.

Each case is a small grid chosen to exercise one part of the lattice update.  The inputs are
regenerated from the seed by lbm_amd.inputs (their SHA-256 is stored to catch generator
drift); turbulence fields are stored in full.  The stored output is the restart record the
reference writes after ``steps`` steps: f1 including ghost layers (src/main.F90:398).
"""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

# name: (q, prec, dims, steps, seed, experiment, turbulence nrturb or 0, options)
CASES = {
    # the BGK box: D3Q19 BGK, periodic i/j, no-slip k, tau = 0.6
    "q19_bgk_pp11": (19, "sp", (16, 8, 6), 20, 1, "none", 0,
                     dict(ibgk=2, ihrr=0, ivreman=0, ibnd=0, jbnd=0, kbnd=11, visc=1.0 / 30.0)),
    # the cylinder wake: D3Q27 3rd-order regularized + Vreman, inflow/outflow, periodic j,
    # no-slip k=1, free-slip k=nz, the reference's cylinder (r=7 at i=50, j=ny/2), turbulence
    # forcing at i=10
    "q27_regvre_inflow_cyl": (27, "sp", (60, 14, 5), 20, 2, "cylinder", 25,
                              dict(ibgk=3, ihrr=1, ivreman=1, smagorinsky=0.16733, ibnd=1, jbnd=0,
                                   kbnd=12, uini=0.05, udir=0.0, visc=1e-4, ampl=1e-3)),
    # the same case over 100 steps: the moments' 1e-5 bar after 100 steps (nrturb above the
    # step count: the reference regenerates the fields with FFTW every nrturb steps,
    # src/main.F90:381-383, and FFTW is absent here, so the injected fields serve throughout)
    "q27_regvre_inflow_cyl_100": (27, "sp", (60, 14, 5), 100, 2, "cylinder", 110,
                                  dict(ibgk=3, ihrr=1, ivreman=1, smagorinsky=0.16733, ibnd=1, jbnd=0,
                                       kbnd=12, uini=0.05, udir=0.0, visc=1e-4, ampl=1e-3)),
    # the same physics in the FP64 build (no cylinder: grid too small for it)
    "q27_regvre_inflow_dp": (27, "dp", (16, 8, 6), 20, 3, "none", 25,
                             dict(ibgk=3, ihrr=1, ivreman=1, smagorinsky=0.16733, ibnd=1, jbnd=0,
                                  kbnd=12, uini=0.05, udir=0.0, visc=1e-4, ampl=1e-3)),
    # all walls closed, mixed no-slip/free-slip: closed-wall kernels, edges and corners
    "q27_regvre_closed": (27, "sp", (12, 8, 6), 20, 5, "none", 0,
                          dict(ibgk=3, ihrr=1, ivreman=1, smagorinsky=0.16733, ibnd=22, jbnd=11,
                               kbnd=12, visc=1e-3)),
    "q19_bgk_closed": (19, "sp", (12, 8, 6), 20, 3, "none", 0,
                       dict(ibgk=2, ibnd=11, jbnd=21, kbnd=22, visc=1.0 / 30.0)),
    # equal dimensions make every entry of the reference's corner table a real corner
    "q27_bgk3_closed_cube": (27, "sp", (8, 8, 8), 12, 7, "none", 0,
                             dict(ibgk=3, ihrr=0, ivreman=0, ibnd=12, jbnd=21, kbnd=11, visc=0.02)),
    # D3Q19 regularized + Vreman, periodic i and k, closed j
    "q19_regvre_pj12": (19, "sp", (12, 8, 6), 20, 4, "none", 0,
                        dict(ibgk=3, ihrr=1, ivreman=1, smagorinsky=0.2, ibnd=0, jbnd=12, kbnd=0,
                             visc=1e-3)),
    # inflow at 20 degrees, free-slip j, no-slip k
    "q19_bgk_inflow_dir": (19, "sp", (12, 8, 6), 20, 6, "none", 0,
                           dict(ibgk=2, ibnd=1, jbnd=22, kbnd=11, uini=0.03, udir=20.0, visc=1.0 / 30.0)),
    # inflow profile and time series read from uvel_shear.dat / uvel_time.dat (FILES below)
    "q27_bgk3_inflow_series": (27, "sp", (16, 8, 6), 20, 8, "none", 0,
                               dict(ibgk=3, ihrr=0, ivreman=0, ibnd=1, jbnd=0, kbnd=11, uini=0.05, udir=0.0,
                                    visc=1.0 / 30.0)),
    # actuator-line turbines (TURBINES below), D3Q27 reg + Vreman, inflow/outflow, one rotor
    "q27_regvre_turbine": (27, "sp", (16, 20, 20), 20, 9, "none", 0,
                           dict(ibgk=3, ihrr=1, ivreman=1, smagorinsky=0.16733, ibnd=1, jbnd=0, kbnd=12,
                                uini=0.05, udir=0.0, visc=0.01)),
    # two yawed and tilted rotors whose deposits overlap, D3Q19 BGK, periodic i and j, no-slip k
    "q19_bgk_turbines2": (19, "sp", (28, 20, 20), 20, 10, "none", 0,
                          dict(ibgk=2, ibnd=0, jbnd=0, kbnd=11, uini=0.04, udir=0.0, visc=0.02)),
    # buoyancy (iablvisc = 2, stable: no surface heat flux), D3Q27 reg + Vreman, periodic i and j,
    # no-slip k=1 / free-slip k=nz; the reference advects its potential temperature every step
    "q27_regvre_buoyancy": (27, "sp", (12, 10, 12), 10, 11, "none", 0,
                            dict(ibgk=3, ihrr=1, ivreman=1, smagorinsky=0.16733, ibnd=0, jbnd=0, kbnd=12,
                                 uini=0.03, udir=0.0, visc=0.01)),
}

# Buoyancy cases: the initial potential temperature 300 K + 0.5 K per cell above k = 6 + U(-2, 2) K
# (seeded), stability class 1 (src/m_heatflux.F90:30-45 adds no flux), in the units of the
# turbine cases (exact conversions); scaling = g * p2l%time^2 / p2l%length.
BUOYANCY = {"q27_regvre_buoyancy": dict(istable=1, ablheight=400.0, theta0=300.0)}


def initial_pottemp(name):
    q, prec, dims, steps, seed, exp, nrturb, o = CASES[name]
    nx, ny, nz = dims
    rng = np.random.default_rng(1000 + seed)
    k = np.arange(nz + 2, dtype=np.float64)[:, None, None]
    th = 300.0 + 0.5 * (k - 6) + rng.uniform(-2.0, 2.0, (nz + 2, ny + 2, nx + 2))
    return th.astype(np.float32 if prec == "sp" else np.float64)

# Turbines of a case (infile section src/m_readinfile.F90:143-161): (ipos, jpos, kpos, yaw, tilt [deg]),
# RPM, pitch [deg], and the azimuths theta [rad] of the restart record.  C_l = 16 m and C_u = 256 m/s
# (rotor radius 64.4 m = 4.03 cells): every conversion of src/m_readinfile.F90:130-186 is then an exact
# power-of-two scaling, so the infile's physical values give the case's lattice uini and kinevisc exactly.
TURBINES = {
    "q27_regvre_turbine": dict(pos=[(6, 10, 10, 0.0, 0.0)], rpm=9.0, pitch=0.0, theta=[0.0]),
    "q19_bgk_turbines2": dict(pos=[(9, 10, 10, 15.0, 5.0), (16, 10, 11, -20.0, 0.0)], rpm=12.0, pitch=2.0,
                              theta=[0.3, 1.1]),
}
TURBINE_LENGTH, TURBINE_VEL = 16.0, 256.0


def infile_units(name):
    """Extra infile_text arguments of a case: physical units, the turbine section, the atmospheric
    boundary layer section."""
    if name in BUOYANCY:
        q, prec, dims, steps, seed, exp, nrturb, o = CASES[name]
        b = BUOYANCY[name]
        return dict(length=TURBINE_LENGTH, vel=TURBINE_VEL, uini=o.get("uini", 0.0) * TURBINE_VEL,
                    visc=o["visc"] * TURBINE_LENGTH * TURBINE_VEL, iablvisc=2, istable=b["istable"],
                    ablheight=b["ablheight"])
    if name not in TURBINES:
        return {}
    q, prec, dims, steps, seed, exp, nrturb, o = CASES[name]
    t = TURBINES[name]
    return dict(length=TURBINE_LENGTH, vel=TURBINE_VEL, uini=o.get("uini", 0.0) * TURBINE_VEL,
                visc=o["visc"] * TURBINE_LENGTH * TURBINE_VEL,
                turbines=dict(pos=t["pos"], rpm=t["rpm"], pitch=t["pitch"], tsr=0.0, tiploss=0))

# Input files of a case, written into the reference's working directory (src/m_inflow.F90:27-86
# reads them): a power-law profile over the nz = 6 cells, and a series of speeds and directions
# whose rows start after step 1 (t = 0), interpolate in between and end before the last step.
FILES = {
    "q27_bgk3_inflow_series": {
        "uvel_shear.dat": "".join(f"{k:4d} {k - 0.5:8.3f} {((k - 0.5) / 6) ** 0.2:.7f}\n" for k in range(1, 7)),
        "uvel_time.dat": "   3.0  0.020   5.0\n   9.5  0.050 -10.0\n  14.0  0.035  30.0\n  17.0  0.040  30.0\n",
    },
}


def case_files(name):
    return FILES.get(name, {})


def case_inputs(name):
    """Initial restart state and turbulence fields of a case (regenerated from its seed)."""
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    from lbm_amd import inputs
    q, prec, dims, steps, seed, exp, nrturb, o = CASES[name]
    dt = np.float32 if prec == "sp" else np.float64
    u0 = (o.get("uini", 0.05), 0.0, 0.0)
    f0 = inputs.perturbed_uniform_flow(q, dims, seed=seed, u0=u0, order=o.get("ibgk", 2), dtype=dt)
    turb = None
    if nrturb:
        turb = inputs.turbulence_fields(seed, dims[1], dims[2], nrturb, dtype=dt)
    return f0, turb


def blanking(name):
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    from lbm_amd import inputs
    q, prec, dims, steps, seed, exp, nrturb, o = CASES[name]
    if exp == "cylinder":
        # the reference's own geometry: radius 7 at (50, nyg/2) (src/m_cylinder.F90:10)
        return inputs.cylinder_mask(dims, 50, dims[1] // 2, 7)
    return None


def fixture_path(name):
    return os.path.join(HERE, f"{name}.npz")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def load(name):
    z = np.load(fixture_path(name), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    out = {k: z[k] for k in z.files if k != "meta"}
    return meta, out
