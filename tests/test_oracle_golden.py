"""The C restatement (oracle/lbm_oracle.c) against the golden vectors the reference's own
build wrote (tests/golden, produced by tests/golden/make_golden.py).  Bit-exact: both evaluate
the same operations in the same order in IEEE FP32/FP64."""
import numpy as np
import pytest

import cases
import oracle


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
    f = sim.run(1, 1 + steps, uini=float(dt(o.get("uini", 0.0))), udir=float(dt(o.get("udir", 0.0))),
                blank=cases.blanking(name), turb=tb)
    ref = fx["f_out"]
    assert f.dtype == ref.dtype and f.shape == ref.shape
    assert np.array_equal(f, ref), f"max diff {np.abs(f - ref).max()}"
