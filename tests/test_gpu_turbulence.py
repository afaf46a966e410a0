"""Device inflow-turbulence generator (lbm_generate_turbulence) against the numpy oracle
(oracle/pseudo2d.py, pinned by tests/test_oracle_turbulence.py) on the same seed: FP64 compute
within 1e-9, FP32 within 2e-4 (the AR(1) recursion and normalization run in the compute type),
for even and odd ny; the generated fields drive the step exactly as the same fields passed
through lbm_set_turbulence."""
import numpy as np
import pytest

import pseudo2d
from lbm_amd import inputs
from lbm_amd.domain import Domain

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage,ny,tol", [("f64", 24, 1e-9), ("f64", 23, 1e-9), ("f32", 24, 2e-4)])
def test_generator_matches_oracle(storage, ny, tol):
    dims, nrturb, seed = (30, ny, 20), 30, 5
    print(f"turbulence {storage} dims={dims} nrturb={nrturb}", flush=True)
    d = Domain(19, dims, storage, ibgk=3, ihrr=1, ivreman=1, smagorinsky=0.16733, kinevisc=1e-3, ibnd=1, kbnd=12)
    got = d.generate_turbulence(nrturb, 1e-3, seed)
    ref = pseudo2d.generate(ny, dims[2], nrturb, seed)
    for g, r in zip(got, ref):
        assert g.shape == r.shape
        err = np.abs(g.astype(np.float64) - r).max()
        print(f"  max |diff| {err:.3e}", flush=True)
        assert err <= tol
    d.close()


def test_generated_fields_drive_the_step_like_set_turbulence():
    dims = (40, 24, 16)
    f0 = inputs.perturbed_uniform_flow(19, dims, seed=9, order=3)
    kw = dict(ibgk=3, ihrr=1, ivreman=1, smagorinsky=0.16733, kinevisc=1e-3, ibnd=1, kbnd=12)
    a, b = Domain(19, dims, "f32", **kw), Domain(19, dims, "f32", **kw)
    uu, vv, ww, rr = a.generate_turbulence(12, 1e-3, 77)
    b.set_turbulence(uu, vv, ww, 12, 1e-3)
    for d in (a, b):
        d.set_inflow(0.05)
        d.set_state(f0)
        d.step(1, 30)
    assert np.array_equal(a.get_state(), b.get_state())
    a.close()
    b.close()
