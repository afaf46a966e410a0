"""Instability detection (src/m_diag.F90:169-173 stops on a negative density): lbm_get_totals and
lbm_get_macros return LBM_EUNSTABLE for a state holding a non-finite population or a cell with
density <= 0, and keep returning the values.  The bad states are restart records with a NaN, or
one cell's populations negated; the kernels only do arithmetic on them (NaN propagates, nothing
faults)."""
import numpy as np
import pytest

from lbm_amd import inputs
from lbm_amd.domain import Domain, LBMError

pytestmark = pytest.mark.gpu
EUNSTABLE = 4


def domain(storage="f32"):
    d = Domain(27, (12, 10, 8), storage, ibgk=3, ihrr=1, ivreman=1, smagorinsky=0.17, kinevisc=0.01, ibnd=1, jbnd=0,
               kbnd=12)
    d.set_inflow(0.05)
    return d


def codes(dom):
    out = []
    for call in (dom.totals, dom.macrovars):
        try:
            call()
            out.append(0)
        except LBMError as e:
            out.append(EUNSTABLE if f"failed ({EUNSTABLE})" in str(e) else -1)
    return out


@pytest.mark.parametrize("storage", ["f32", "f64", "f16s"])
def test_stable_state_passes(storage):
    dt = np.float64 if storage == "f64" else np.float32
    f0 = inputs.perturbed_uniform_flow(27, (12, 10, 8), seed=4, u0=(0.05, 0, 0), order=3, dtype=dt)
    with domain(storage) as d:
        print(f"stable {storage}", flush=True)
        d.set_state(f0)
        assert codes(d) == [0, 0]
        d.step(1, 5)
        assert codes(d) == [0, 0]


@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_nan_population_is_unstable(storage):
    dt = np.float64 if storage == "f64" else np.float32
    f0 = inputs.perturbed_uniform_flow(27, (12, 10, 8), seed=4, u0=(0.05, 0, 0), order=3, dtype=dt)
    f0[4, 5, 6, 7] = np.nan          # population 7 of cell (6, 5, 4)
    with domain(storage) as d:
        print(f"nan {storage}", flush=True)
        d.set_state(f0)
        assert codes(d) == [EUNSTABLE, EUNSTABLE]
        with pytest.raises(LBMError, match="unstable state: 1 non-finite populations"):
            d.totals()
        d.step(1, 3)                 # the NaN spreads through the collisions; still reported
        assert codes(d) == [EUNSTABLE, EUNSTABLE]


def test_negative_density_is_unstable():
    f0 = inputs.perturbed_uniform_flow(27, (12, 10, 8), seed=4, u0=(0.05, 0, 0), order=3)
    f0[3, 4, 5] *= -1.0              # cell (5, 4, 3): density -rho
    with domain() as d:
        print("negative density", flush=True)
        d.set_state(f0)
        assert codes(d) == [EUNSTABLE, EUNSTABLE]
        with pytest.raises(LBMError, match="smallest cell density -"):
            d.totals()
