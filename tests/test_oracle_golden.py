"""The C restatement (oracle/lbm_oracle.c) against the golden vectors the reference's own
build wrote (tests/golden, produced by tests/golden/make_golden.py).  Bit-exact: both evaluate
the same operations in the same order in IEEE FP32/FP64."""
import numpy as np
import pytest

import cases
import oracle


def turbines_of(name, fx):
    """The oracle's turbine description of a case (turbine_initialize on its infile values), with
    the airfoil tables stored in the fixture."""
    if name not in cases.TURBINES:
        return None
    t = cases.TURBINES[name]
    prec = "f32" if cases.CASES[name][1] == "sp" else "f64"
    foils = [fx["foil"][k, :n] for k, n in enumerate(fx["nang"])]
    return oracle.turbines_from_infile(t["pos"], t["rpm"], t["pitch"], cases.TURBINE_LENGTH, cases.TURBINE_VEL,
                                       foils, prec, theta=t["theta"])


def buoyancy_of(name, fx):
    """(potential temperature of every step, scaling, theta0) of a buoyancy case: the fields the
    reference's own transport produced (stored in the fixture)."""
    if name not in cases.BUOYANCY:
        return None
    from lbm_amd.domain import buoyancy_scaling
    prec = cases.CASES[name][1]
    dt = np.float32 if prec == "sp" else np.float64
    return (fx["pottemp"], buoyancy_scaling(cases.TURBINE_LENGTH, cases.TURBINE_VEL, dt),
            float(dt(cases.BUOYANCY[name]["theta0"])))


@pytest.mark.parametrize("name", sorted(cases.CASES))
def test_oracle_matches_reference_build(name):
    meta, fx = cases.load(name)
    q, prec, dims, steps, seed, exp, nrturb, o = cases.CASES[name]
    f0, turb = cases.case_inputs(name)
    assert cases.sha(f0) == meta["input_sha256"], "seeded input generator changed"
    dt = np.float32 if prec == "sp" else np.float64
    sim = oracle.OracleSim(q, dims, "f32" if prec == "sp" else "f64", ibgk=o.get("ibgk", 2),
                           ihrr=o.get("ihrr", 0), ivreman=o.get("ivreman", 0),
                           smagorinsky=float(dt(o.get("smagorinsky", 0.0))),
                           kinevisc=float(dt(o.get("visc", 1 / 30.0))),
                           ibnd=o["ibnd"], jbnd=o["jbnd"], kbnd=o["kbnd"])
    sim.restart(f0)
    tb = None
    if nrturb:
        tb = (fx["uu"], fx["vv"], fx["ww"], nrturb, float(dt(o.get("ampl", 0.0))))
    uini, udir = float(dt(o.get("uini", 0.0))), float(dt(o.get("udir", 0.0)))
    shear, ut, ud = oracle.inflow_series(cases.case_files(name), dims[2], 1, 1 + steps, uini, udir,
                                         "f32" if prec == "sp" else "f64")
    T = turbines_of(name, fx)
    B = buoyancy_of(name, fx)
    f = sim.run(1, 1 + steps, uvel_shear=shear, uvel_time=ut, udir_time=ud, blank=cases.blanking(name), turb=tb,
                turbines=T, buoyancy=B)
    ref = fx["f_out"]
    if T is not None:
        assert np.array_equal(T["theta"], fx["theta_out"])
    assert f.dtype == ref.dtype and f.shape == ref.shape
    assert np.array_equal(f, ref), f"max diff {np.abs(f - ref).max()}"


def test_inflow_files_drive_the_golden_case(tmp_path):
    """The profile and series of q27_bgk3_inflow_series reach the result: the same run with the
    constant inflow of infile.in does not reproduce the golden record.  The product's reader of
    the two files (lbm_amd.fortran_io.read_inflow) gives the oracle's values bit for bit."""
    from lbm_amd import fortran_io
    name = "q27_bgk3_inflow_series"
    meta, fx = cases.load(name)
    q, prec, dims, steps, seed, exp, nrturb, o = cases.CASES[name]
    files = cases.case_files(name)
    assert meta["files"] == files
    f0, _ = cases.case_inputs(name)
    shear, ut, ud = oracle.inflow_series(files, dims[2], 1, 1 + steps, o["uini"], o["udir"])
    assert len(set(ut[1:].tolist())) > 5 and len(set(ud[1:].tolist())) > 5 and shear[-1] == 1 and shear[0] < 0.8
    for fname, text in files.items():
        (tmp_path / fname).write_text(text)
    ps, pt, pd = fortran_io.read_inflow(str(tmp_path), dims[2], 1, 1 + steps, o["uini"], o["udir"])
    assert np.array_equal(ps, shear) and np.array_equal(pt, ut[1:]) and np.array_equal(pd, ud[1:])
    sim = oracle.OracleSim(q, dims, "f32", ibgk=3, kinevisc=float(np.float32(o["visc"])), ibnd=1, jbnd=0, kbnd=11)
    sim.restart(f0)
    f = sim.run(1, 1 + steps, uini=float(np.float32(o["uini"])))
    assert np.abs(f - fx["f_out"]).max() > 1e-4


@pytest.mark.parametrize("name", sorted(cases.TURBINES))
def test_turbines_drive_the_golden_cases(name):
    """Without its turbines a turbine case's run does not reproduce the golden record: the golden
    pins the turbine forcing and forcings_apply (src/main.F90:291-306), whose increment there is
    several orders above round-off."""
    meta, fx = cases.load(name)
    q, prec, dims, steps, seed, exp, nrturb, o = cases.CASES[name]
    assert meta["turbines"]["pos"] == [list(p) for p in cases.TURBINES[name]["pos"]]
    f0, _ = cases.case_inputs(name)
    sim = oracle.OracleSim(q, dims, "f32", ibgk=o.get("ibgk", 2), ihrr=o.get("ihrr", 0), ivreman=o.get("ivreman", 0),
                           smagorinsky=float(np.float32(o.get("smagorinsky", 0.0))),
                           kinevisc=float(np.float32(o["visc"])), ibnd=o["ibnd"], jbnd=o["jbnd"], kbnd=o["kbnd"])
    sim.restart(f0)
    f = sim.run(1, 1 + steps, uini=float(np.float32(o.get("uini", 0.0))))
    assert np.abs(f - fx["f_out"]).max() > 1e-4   # measured 2.8e-4 and 3.4e-4
    assert fx["theta_out"].shape == (len(cases.TURBINES[name]["pos"]),)


@pytest.mark.parametrize("name", sorted(cases.BUOYANCY))
def test_buoyancy_drives_the_golden_cases(name):
    """Without its buoyancy a buoyancy case's run does not reproduce the golden record (the golden
    pins buoyancy_forcing and forcings_apply, src/main.F90:293-306); the stored fields are the
    reference's: the first is the seeded initial one and its transport changes them every step."""
    meta, fx = cases.load(name)
    q, prec, dims, steps, seed, exp, nrturb, o = cases.CASES[name]
    assert np.array_equal(fx["pottemp"][0], cases.initial_pottemp(name))
    assert fx["pottemp"].shape == (steps, dims[2] + 2, dims[1] + 2, dims[0] + 2)
    assert all(np.abs(fx["pottemp"][n + 1] - fx["pottemp"][n]).max() > 1e-3 for n in range(steps - 1))
    f0, _ = cases.case_inputs(name)
    sim = oracle.OracleSim(q, dims, "f32", ibgk=o.get("ibgk", 2), ihrr=o.get("ihrr", 0), ivreman=o.get("ivreman", 0),
                           smagorinsky=float(np.float32(o.get("smagorinsky", 0.0))),
                           kinevisc=float(np.float32(o["visc"])), ibnd=o["ibnd"], jbnd=o["jbnd"], kbnd=o["kbnd"])
    sim.restart(f0)
    f = sim.run(1, 1 + steps, uini=float(np.float32(o.get("uini", 0.0))))
    assert np.abs(f - fx["f_out"]).max() > 1e-5   # measured 6.6e-5
    # the first step's field alone instead of the transported ones: also off the record
    sim.restart(f0)
    B = buoyancy_of(name, fx)
    f1 = sim.run(1, 1 + steps, uini=float(np.float32(o.get("uini", 0.0))),
                 buoyancy=(np.repeat(fx["pottemp"][:1], steps, axis=0), B[1], B[2]))
    assert not np.array_equal(f1, fx["f_out"])
