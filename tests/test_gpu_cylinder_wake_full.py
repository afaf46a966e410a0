"""The headline workload at its stated size against the reference's own build: D3Q27 third-order
regularized + Vreman (c = 0.07), 384x192x128, inflow u = 0.05 at i = 1 / outflow at i = nx,
periodic j, no-slip k = 1, free-slip k = nz, solid cylinder r = 12 on the k axis at (96, 96),
seeded inflow forcing at i = 10 (amplitude 1e-3), nu = 1e-4.  The reference executable is its
OpenMP build for exactly this grid and cylinder (oracle/build_ref.py; the CPU baseline of
bench.py runs it too); every cell's update is independent of the thread count.  Both start from the
same restart record and turbulence fields and take 20 steps; the library's state must match the
reference's restart record at the golden bar (populations 2e-6, moments 1e-5 relative).  Skipped
where the reference build is absent."""
import os
import tempfile

import numpy as np
import pytest

import build_ref
import oracle
from lbm_amd import inputs
from lbm_amd.domain import Domain
from lbm_amd.lattice import lattice

pytestmark = pytest.mark.gpu

DIMS, STEPS, SEED = (384, 192, 128), 20, 2
SMAG = float(np.sqrt(0.07 / 2.5))


def rel_moment_errors(f, ref, q):
    lat = lattice(q)

    def mom(a):
        a = a[1:-1, 1:-1, 1:-1].astype(np.float64)
        r = a.sum(-1)
        return r, (a * lat.cx).sum(-1) / r, (a * lat.cy).sum(-1) / r, (a * lat.cz).sum(-1) / r

    r1, u1, v1, w1 = mom(f)
    r0, u0, v0, w0 = mom(ref)
    speed = max(np.abs(u0).max(), np.abs(v0).max(), np.abs(w0).max())
    return (np.abs(r1 - r0).max() / np.abs(r0).max(),
            max(np.abs(u1 - u0).max(), np.abs(v1 - v0).max(), np.abs(w1 - w0).max()) / speed)


def test_headline_workload_full_size_matches_reference_build():
    exe = build_ref.exe_path(27, "sp", DIMS, omp=True, cylinder=(96, 96, 12))
    if not os.path.exists(exe):
        pytest.skip("the reference's OpenMP build for 384x192x128 is not built here")
    nx, ny, nz = DIMS
    nr = STEPS + 5
    f0 = inputs.perturbed_uniform_flow(27, DIMS, seed=SEED, u0=(0.05, 0, 0), order=3)
    turb = inputs.turbulence_fields(SEED, ny, nz, nr, corr=8.0)
    mask = inputs.cylinder_mask(DIMS, 96, 96, 12)
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    print(f"reference build {DIMS} {STEPS} steps on {threads} threads", flush=True)
    with tempfile.TemporaryDirectory(prefix="lbm_cfg2_") as wd:
        ref, _ = oracle.run_reference(exe, f0, nt0=1, nt1=1 + STEPS, turb_fields=turb, workdir=wd, threads=threads,
                                      ibgk=3, ihrr=1, ivreman=1, smagorinsky=float(np.float32(SMAG)), ibnd=1, jbnd=0,
                                      kbnd=12, uini=0.05, udir=0.0, visc=1e-4, ampl=1e-3, nrturb=nr,
                                      experiment="cylinder", ltiming=False, ltesting=False, ldump=True)
    print("library", flush=True)
    with Domain(27, DIMS, "f32", ibgk=3, ihrr=1, ivreman=1, smagorinsky=SMAG, kinevisc=1e-4, ibnd=1, jbnd=0,
                kbnd=12) as dom:
        dom.set_state(f0)
        dom.set_inflow(0.05)
        dom.set_blanking(mask)
        dom.set_turbulence(turb[0], turb[1], turb[2], nr, 1e-3)
        dom.step(1, STEPS)
        f = dom.get_state()
    assert f.shape == ref.shape and f.dtype == ref.dtype
    assert np.isfinite(f).all()
    d = np.abs(f.astype(np.float64) - ref.astype(np.float64))
    er, eu = rel_moment_errors(f, ref, 27)
    print(f"max |df| {d.max():.3e}, rho {er:.3e}, velocity {eu:.3e}", flush=True)
    assert d.max() <= 2e-6, f"max |df| {d.max():.3e} at {np.unravel_index(d.argmax(), d.shape)}"
    assert er <= 1e-5 and eu <= 1e-5, (er, eu)
