"""lbm_get_totals_noalloc (the totals without any device allocation: partial results in the
library's fixed per-GPU scratch array) against lbm_get_totals on the same states, through the C
ABI: bit-identical mass and momentum, two calls in a row with the second on a smaller grid (so a
larger earlier call's partial results are left in the scratch), the largest grid large enough for
the totals kernel to grid-stride (more than 1024 x 256 cells), and the same LBM_EUNSTABLE status
on a state holding a NaN."""
import numpy as np
import pytest

from lbm_amd import inputs
from lbm_amd.domain import Domain, LBMError

pytestmark = pytest.mark.gpu


def stepped(dims, seed, steps=3):
    f0 = inputs.perturbed_uniform_flow(19, dims, seed=seed, u0=(0.05, 0, 0), order=2)
    d = Domain(19, dims, "f32", ibgk=2, ihrr=0, ivreman=0, kinevisc=0.02, ibnd=0, jbnd=0, kbnd=11)
    d.set_state(f0)
    d.step(1, steps)
    return d


def test_noalloc_totals_match_two_calls_shrinking():
    big, small = (96, 64, 48), (12, 10, 8)      # 294912 cells (1024 blocks, grid-stride), 960 cells
    print(f"no-alloc totals {big} then {small}", flush=True)
    with stepped(big, 5) as db, stepped(small, 6) as ds:
        m_ref, p_ref = db.totals()
        m_na, p_na = db.totals(noalloc=True)
        assert m_na == m_ref and np.array_equal(p_na, p_ref)
        assert abs(m_na - np.prod(big)) < 1e-3 * np.prod(big)
        # second call: a smaller grid, after the larger one filled more of the scratch
        m_ref2, p_ref2 = ds.totals()
        m_na2, p_na2 = ds.totals(noalloc=True)
        assert m_na2 == m_ref2 and np.array_equal(p_na2, p_ref2)
        assert abs(m_na2 - np.prod(small)) < 1e-3 * np.prod(small)
        # and back: the first context again after a step
        db.step(4, 1)
        m_na3, p_na3 = db.totals(noalloc=True)
        m_ref3, p_ref3 = db.totals()
        assert m_na3 == m_ref3 and np.array_equal(p_na3, p_ref3)


def test_noalloc_totals_unstable_status():
    dims = (12, 10, 8)
    f0 = inputs.perturbed_uniform_flow(19, dims, seed=7, u0=(0.05, 0, 0), order=2)
    f0[3, 4, 5, 2] = np.nan
    print("no-alloc totals, NaN state", flush=True)
    d = Domain(19, dims, "f32", ibgk=2, ihrr=0, ivreman=0, kinevisc=0.02, ibnd=0, jbnd=0, kbnd=11)
    with d:
        d.set_state(f0)
        for noalloc in (False, True):
            with pytest.raises(LBMError, match="unstable state: 1 non-finite populations"):
                d.totals(noalloc=noalloc)
