"""Low-precision population storage (weight-shifted FP16 and BF16, FP32 arithmetic) and the step
kernel's form for population buffers of 4 GiB and more.

Error measure (as for the FP16/BF16 targets): relative L2 of the velocity field against the
library's FP64 storage run on the same input, ||u - u64||_2 / ||u64||_2 over interior cells (u the
3-vector), and the same for rho - 1.  The FP64 path is pinned to the FP64 oracle by
test_gpu_oracle.py::test_cylinder_wake_physics[f64] at 1e-12.
"""
import json
import os

import numpy as np
import pytest

import oracle
from lbm_amd import inputs
from lbm_amd.domain import Domain
from lbm_amd.lattice import lattice

pytestmark = pytest.mark.gpu

WAKE = dict(ibgk=3, ihrr=1, ivreman=1, smagorinsky=0.16733, kinevisc=1e-4, ibnd=1, jbnd=0, kbnd=12)


def run(storage, dims, steps, f0, tu, mask, env=None):
    """env: variables set while the context is created (the library reads them in lbm_create)."""
    env = env or {}
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        dom = Domain(27, dims, storage, **WAKE)
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v
    dom.set_state(f0)
    dom.set_inflow(0.05)
    dom.set_blanking(mask)
    dom.set_turbulence(tu[0], tu[1], tu[2], len(tu[0]) - 1, 1e-3)
    dom.step(1, steps)
    r, u, v, w = (a[1:-1, 1:-1, 1:-1].astype(np.float64) for a in dom.macrovars(np.float64))
    f = dom.get_state(np.float64)
    dom.close()
    return f, r, np.stack([u, v, w])


def rel_l2(a, b):
    return float(np.sqrt(((a - b) ** 2).sum()) / np.sqrt((b ** 2).sum()))


@pytest.fixture(scope="module")
def case():
    dims = (96, 32, 16)
    f0 = inputs.perturbed_uniform_flow(27, dims, seed=21, order=3, dtype=np.float64)
    tu = inputs.turbulence_fields(21, dims[1], dims[2], 120, dtype=np.float64)
    mask = inputs.cylinder_mask(dims, 40, 16, 6)
    return dims, f0, tu, mask


@pytest.mark.parametrize("storage,tol", [("f16s", 1e-3), ("bf16s", 1e-2)])
def test_lowprec_error_vs_fp64(case, storage, tol):
    dims, f0, tu, mask = case
    steps = 100
    print(f"{storage} {dims} {steps}", flush=True)
    _, r64, u64 = run("f64", dims, steps, f0, tu, mask)
    _, r32, u32 = run("f32", dims, steps, f0, tu, mask)
    _, rlo, ulo = run(storage, dims, steps, f0, tu, mask)
    e_lo, e_32 = rel_l2(ulo, u64), rel_l2(u32, u64)
    er_lo = rel_l2(rlo - 1.0, r64 - 1.0)
    print(f"rel-L2 velocity: {storage} {e_lo:.3e}  f32 {e_32:.3e}; rho-1: {er_lo:.3e}", flush=True)
    assert np.isfinite(ulo).all()
    assert e_lo <= tol
    assert e_32 <= 1e-5


# The forms a population buffer of 4 GiB or more takes (D3Q27 FP32 at 768x256x192 is 5.2 GB): the
# row-blocked arrangement addressed per slab of the slowest axis (the default above 4 GiB while a
# slab stays below it), and the plane arrangement with 64-bit plane offsets.  Variant bit 1 (and
# LBM_LAYOUT=planes) force them on a small grid.
SLAB = {"LBM_STEP_VARIANT": "3"}
LARGE = {"LBM_LAYOUT": "planes", "LBM_STEP_VARIANT": "3"}


@pytest.mark.parametrize("storage", ["f32", "f16s", "bf16s", "f64"])
def test_large_buffer_form_matches_default(case, storage):
    """The >= 4 GiB forms give the default form's states bit for bit (the same arithmetic, other
    addressing), after steps through the folded face, the inflow/outflow planes, the cylinder and
    the forcing slice."""
    dims, f0, tu, mask = case
    steps = 30
    if storage != "f64":
        f0 = f0.astype(np.float32)
        tu = tuple(a.astype(np.float32) for a in tu)
    print(f"large-buffer forms {storage} {dims} {steps} steps", flush=True)
    fa, _, _ = run(storage, dims, steps, f0, tu, mask)
    for env in (SLAB, LARGE):
        fb, _, _ = run(storage, dims, steps, f0, tu, mask, env=env)
        assert np.isfinite(fb).all()
        assert np.array_equal(fa, fb), (env, float(np.abs(fa - fb).max()))


def test_large_buffer_form_fp32_matches_oracle(case):
    """The >= 4 GiB form against the oracle (FP32)."""
    dims, f0, tu, mask = case
    steps = 60
    f0 = f0.astype(np.float32)
    tu32 = tuple(a.astype(np.float32) for a in tu)
    f, _, _ = run("f32", dims, steps, f0, tu32, mask, env=LARGE)
    so = {k: (float(np.float32(v)) if isinstance(v, float) else v) for k, v in WAKE.items()}
    sim = oracle.OracleSim(27, dims, "f32", **so)
    sim.restart(f0)
    sim.run(1, 1 + steps, uini=float(np.float32(0.05)), blank=mask,
            turb=(tu32[0], tu32[1], tu32[2], len(tu32[0]) - 1, float(np.float32(1e-3))))
    d = np.abs(f - sim.f1.astype(np.float64))
    print(f"large-buffer form fp32 vs oracle: max|df| {d.max():.3e}", flush=True)
    assert d.max() <= 1e-5


@pytest.mark.parametrize("storage", ["f16s", "bf16s"])
def test_lowprec_roundtrip_and_odd_length(storage):
    """set_state/get_state round trip within the storage quantum, and a grid whose contiguous
    axis has odd length (one cell per lane) stepping like an even one."""
    q, dims = 19, (15, 9, 7)
    lat = lattice(q)
    f0 = inputs.perturbed_uniform_flow(q, dims, seed=4, order=2)
    dom = Domain(q, dims, storage, ibgk=2, kinevisc=1.0 / 30.0, ibnd=0, jbnd=0, kbnd=11)
    dom.set_state(f0)
    back = dom.get_state()
    dev = np.abs(f0 - lat.w).max()
    quantum = dev * (2.0 ** -11 if storage == "f16s" else 2.0 ** -8)
    assert np.abs(back - f0).max() <= quantum * 1.01 + 1e-8
    dom.step(1, 30)
    f = dom.get_state()
    sim = oracle.OracleSim(q, dims, "f32", ibgk=2, kinevisc=float(np.float32(1 / 30.0)), kbnd=11)
    sim.restart(f0)
    sim.run(1, 31)
    err = np.abs(f - sim.f1).max()
    print(f"{storage} odd-length grid, 30 steps: max|df| vs FP32 oracle {err:.3e}", flush=True)
    assert np.isfinite(f).all() and err <= (5e-4 if storage == "f16s" else 5e-3)


def test_precision_study_runs_small():
    """lbm_amd.precision_study end to end on a small grid (both physics)."""
    from lbm_amd import precision_study
    for phys in ("bgk", "regvre"):
        r = precision_study.main(["--physics", phys, "--dims", "48", "32", "24", "--steps", "10", "30",
                                  "--storages", "f32", "f16s", "--nrturb", "20"])
        e = r["errors"]
        print(phys, json.dumps(e), flush=True)
        assert r["finite"]["f32"] and r["finite"]["f16s"]
        assert e["f32"]["30"]["velocity"] < 1e-5
        assert e["f16s"]["30"]["velocity"] < 1e-2
