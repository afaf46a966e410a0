"""On-device time averaging (lbm_set_averaging / lbm_get_averages) against the same statistics
accumulated on the host from lbm_get_macros after every step (whose moments are pinned to the
oracle): means within 1e-6 (FP32) / 1e-12 (FP64) relative to the largest mean speed, Ti within
2e-4 absolute (FP32; it is a difference of FP32 sums, as in the reference) / 1e-10 (FP64)."""
import numpy as np
import pytest

from lbm_amd import inputs
from lbm_amd.domain import Domain

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage,tol_m,tol_t", [("f32", 1e-6, 2e-4), ("f64", 1e-12, 1e-10)])
def test_averages_match_host_accumulation(storage, tol_m, tol_t):
    dims = (48, 24, 16)
    kw = dict(ibgk=3, ihrr=1, ivreman=1, smagorinsky=0.16733, kinevisc=1e-4, ibnd=1, jbnd=0, kbnd=12)
    d = Domain(27, dims, storage, **kw)
    dt = np.float64 if storage == "f64" else np.float32
    d.set_state(inputs.perturbed_uniform_flow(27, dims, seed=12, order=3, dtype=dt))
    d.set_inflow(0.05)
    d.set_blanking(inputs.cylinder_mask(dims, 16, 12, 4))
    tu = inputs.turbulence_fields(12, dims[1], dims[2], 40, dtype=dt)
    d.set_turbulence(tu[0], tu[1], tu[2], 40, 1e-2)
    d.step(1, 5)
    d.set_averaging(True)
    n = 20
    acc = np.zeros((6,) + tuple(reversed(dims)))
    for s in range(n):
        d.step(6 + s, 1)
        _, u, v, w = (a[1:-1, 1:-1, 1:-1].astype(np.float64) for a in d.macrovars(np.float64))
        acc += np.stack([u, v, w, u * u, v * v, w * w])
    ua, va, wa, ti, count = (x if isinstance(x, int) else x[1:-1, 1:-1, 1:-1].astype(np.float64) for x in d.averages())
    assert count == n
    m = acc[:3] / n
    var = np.maximum(acc[3:] / n - m ** 2, 0.0)
    ti_host = np.sqrt(var.sum(0) / 3.0)
    scale = np.abs(m).max()
    err_m = max(np.abs(ua - m[0]).max(), np.abs(va - m[1]).max(), np.abs(wa - m[2]).max()) / scale
    err_t = np.abs(ti - ti_host).max()
    print(f"{storage}: means rel err {err_m:.2e}, Ti abs err {err_t:.2e} (max Ti {ti_host.max():.2e})", flush=True)
    assert err_m <= tol_m and err_t <= tol_t
    # averaging off: further steps leave the averages alone
    d.set_averaging(False)
    d.step(6 + n, 3)
    assert d.averages()[4] == n
    d.close()
