"""HIP library (through its C-ABI) against the golden vectors of the reference build.

Tolerance (FP32): the library evaluates the collision in closed form and with fused
multiply-adds, so its rounding differs from the reference's dense Hermite contractions.  Over
the 12-100 steps of a fixture the populations must agree to 2e-6 absolute (f is O(0.3)) and the
moments to a relative 1e-5 (density by max|rho|, velocities by the largest speed).  FP64: 1e-12.
"""
import os

import numpy as np
import pytest

import cases
from lbm_amd.lattice import lattice

pytestmark = pytest.mark.gpu


def moments(f, q):
    lat = lattice(q)
    f = f.astype(np.float64)
    rho = f.sum(-1)
    u = (f * lat.cx).sum(-1) / rho
    v = (f * lat.cy).sum(-1) / rho
    w = (f * lat.cz).sum(-1) / rho
    return rho, u, v, w


def rel_errors(f, ref, q):
    r1, u1, v1, w1 = (a[1:-1, 1:-1, 1:-1] for a in moments(f, q))
    r0, u0, v0, w0 = (a[1:-1, 1:-1, 1:-1] for a in moments(ref, q))
    speed = max(np.abs(u0).max(), np.abs(v0).max(), np.abs(w0).max())
    return (np.abs(r1 - r0).max() / np.abs(r0).max(),
            max(np.abs(u1 - u0).max(), np.abs(v1 - v0).max(), np.abs(w1 - w0).max()) / speed)


def run_case(name):
    import tempfile
    from lbm_amd import fortran_io
    from lbm_amd.domain import Domain
    meta, fx = cases.load(name)
    q, prec, dims, steps, seed, exp, nrturb, o = cases.CASES[name]
    f0, _ = cases.case_inputs(name)
    print(f"case {name} q={q} prec={prec} dims={dims} steps={steps}", flush=True)
    dom = Domain(q, dims, "f32" if prec == "sp" else "f64", ibgk=o.get("ibgk", 2), ihrr=o.get("ihrr", 0),
                 ivreman=o.get("ivreman", 0), smagorinsky=o.get("smagorinsky", 0.0),
                 kinevisc=o.get("visc", 1 / 30.0), ibnd=o["ibnd"], jbnd=o["jbnd"], kbnd=o["kbnd"])
    dom.set_state(f0)
    # the inflow files of the case read by the library's reader of src/m_inflow.F90
    with tempfile.TemporaryDirectory() as wd:
        for fname, text in cases.case_files(name).items():
            with open(os.path.join(wd, fname), "w") as fh:
                fh.write(text)
        dt = np.float32 if prec == "sp" else np.float64
        shear, ut, ud = fortran_io.read_inflow(wd, dims[2], 1, 1 + steps, o.get("uini", 0.0), o.get("udir", 0.0),
                                               dtype=dt)
    dom.set_inflow(o.get("uini", 0.0), o.get("udir", 0.0), shear=shear)
    dom.set_inflow_series(1, ut, ud)
    bl = cases.blanking(name)
    if bl is not None:
        dom.set_blanking(bl)
    if nrturb:
        dom.set_turbulence(fx["uu"], fx["vv"], fx["ww"], nrturb, o.get("ampl", 0.0))
    theta = None
    if name in cases.TURBINES:
        from lbm_amd import turbines
        t = cases.TURBINES[name]
        foils = [fx["foil"][k, :n] for k, n in enumerate(fx["nang"])]
        dom.set_turbines(turbines.from_infile(t["pos"], t["rpm"], t["pitch"], cases.TURBINE_LENGTH, cases.TURBINE_VEL,
                                              foils, dt, theta=t["theta"]))
    if name in cases.BUOYANCY:
        # the reference's potential temperature of each step, set before that step
        from lbm_amd.domain import buoyancy_scaling
        sc = buoyancy_scaling(cases.TURBINE_LENGTH, cases.TURBINE_VEL, dt)
        for n in range(steps):
            dom.set_buoyancy(fx["pottemp"][n], theta0=cases.BUOYANCY[name]["theta0"], scaling=sc)
            dom.step(1 + n, 1)
    else:
        dom.step(1, steps)
    f = dom.get_state()
    if name in cases.TURBINES:
        theta = dom.turbine_theta()
        # the azimuths accumulate omegand in the REAL kind, as the reference's
        assert np.array_equal(theta.astype(dt), fx["theta_out"]), (theta, fx["theta_out"])
    dom.close()
    return f, fx["f_out"]


@pytest.mark.parametrize("name", sorted(cases.CASES))
def test_hip_matches_reference_build(name):
    f, ref = run_case(name)
    q = cases.CASES[name][0]
    assert f.shape == ref.shape and f.dtype == ref.dtype
    tol_f = 2e-6 if ref.dtype == np.float32 else 1e-12
    tol_m = 1e-5 if ref.dtype == np.float32 else 1e-12
    d = np.abs(f.astype(np.float64) - ref.astype(np.float64))
    assert np.isfinite(f).all()
    assert d.max() <= tol_f, f"max |df| {d.max():.3e} at {np.unravel_index(d.argmax(), d.shape)}"
    er, eu = rel_errors(f, ref, q)
    assert er <= tol_m and eu <= tol_m, (er, eu)


def test_turbines_and_buoyancy_together_match_oracle():
    """Both external-force producers at once (src/main.F90:290-293: external_forcing = 0, the turbines'
    deposit, then the buoyancy of every interior cell, applied by forcings_apply): the library against
    the oracle on the two-rotor turbine case with a seeded potential temperature per step.  The
    buoyancy moves the result off the turbine-only golden record by far more than the bar."""
    import oracle
    from lbm_amd import turbines
    from lbm_amd.domain import Domain, buoyancy_scaling
    name = "q19_bgk_turbines2"
    meta, fx = cases.load(name)
    q, prec, dims, steps, seed, exp, nrturb, o = cases.CASES[name]
    nx, ny, nz = dims
    f0, _ = cases.case_inputs(name)
    t = cases.TURBINES[name]
    foils = [fx["foil"][k, :n] for k, n in enumerate(fx["nang"])]
    rng = np.random.default_rng(77)
    pot = (300.0 + rng.uniform(-3.0, 3.0, (steps, nz + 2, ny + 2, nx + 2))).astype(np.float32)
    sc = buoyancy_scaling(cases.TURBINE_LENGTH, cases.TURBINE_VEL)
    print(f"turbines + buoyancy {name} {dims} {steps}", flush=True)
    dom = Domain(q, dims, "f32", ibgk=o.get("ibgk", 2), ihrr=o.get("ihrr", 0), ivreman=o.get("ivreman", 0),
                 smagorinsky=o.get("smagorinsky", 0.0), kinevisc=o["visc"], ibnd=o["ibnd"], jbnd=o["jbnd"],
                 kbnd=o["kbnd"])
    dom.set_state(f0)
    dom.set_inflow(o.get("uini", 0.0), o.get("udir", 0.0))
    dom.set_turbines(turbines.from_infile(t["pos"], t["rpm"], t["pitch"], cases.TURBINE_LENGTH, cases.TURBINE_VEL,
                                          foils, np.float32, theta=t["theta"]))
    for n in range(steps):
        dom.set_buoyancy(pot[n], theta0=300.0, scaling=sc)
        dom.step(1 + n, 1)
    f = dom.get_state()
    dom.close()
    sim = oracle.OracleSim(q, dims, "f32", ibgk=o.get("ibgk", 2), ihrr=o.get("ihrr", 0), ivreman=o.get("ivreman", 0),
                           smagorinsky=float(np.float32(o.get("smagorinsky", 0.0))),
                           kinevisc=float(np.float32(o["visc"])), ibnd=o["ibnd"], jbnd=o["jbnd"], kbnd=o["kbnd"])
    sim.restart(f0)
    T = oracle.turbines_from_infile(t["pos"], t["rpm"], t["pitch"], cases.TURBINE_LENGTH, cases.TURBINE_VEL,
                                    foils, "f32", theta=t["theta"])
    ref = sim.run(1, 1 + steps, uini=float(np.float32(o.get("uini", 0.0))), turbines=T, buoyancy=(pot, sc, 300.0))
    d = np.abs(f.astype(np.float64) - ref.astype(np.float64)).max()
    off = np.abs(ref.astype(np.float64) - fx["f_out"].astype(np.float64)).max()
    print(f"max |f - oracle| {d:.3e}; oracle with buoyancy vs turbine-only golden {off:.3e}", flush=True)
    assert np.isfinite(f).all() and d <= 2e-6
    er, eu = rel_errors(f, ref, q)
    assert er <= 1e-5 and eu <= 1e-5, (er, eu)
    assert off > 1e-5
