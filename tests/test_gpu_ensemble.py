"""Ensembles (lbm_ensemble_*): the batched persistent kernels with graph-captured steps must give
every member exactly the state lbm_step gives it on its own context (the same per-cell code
and arithmetic, so bit-identical), for every launch shape, across repeated graph replays with
host work in between, and the autotuner must leave the members' states unchanged.  The
single-domain path itself is pinned to the oracle and the reference build by
test_gpu_golden.py / test_gpu_oracle.py.
"""
import numpy as np
import pytest

from lbm_amd import inputs
from lbm_amd.domain import Domain, Ensemble, LBMError

pytestmark = pytest.mark.gpu

PHYS = dict(ibgk=3, ihrr=1, ivreman=1, smagorinsky=0.16733, kinevisc=1e-3)

# (dims, boundary codes, extras): inflow + closed k with a box and turbulence; closed k only
# with solids; fully periodic with a body force; inflow at 20 degrees with periodic k; closed j
MEMBERS = [
    ((40, 24, 16), dict(ibnd=1, jbnd=0, kbnd=12), dict(box=((10, 8, 1), (15, 13, 6)), turb=True)),
    ((32, 16, 20), dict(ibnd=0, jbnd=0, kbnd=11), dict(box=((5, 5, 5), (9, 9, 9)))),
    ((24, 20, 12), dict(ibnd=0, jbnd=0, kbnd=0), dict(force=True)),
    ((36, 28, 10), dict(ibnd=1, jbnd=0, kbnd=0), dict(udir=20.0, turb=True)),
    ((20, 18, 14), dict(ibnd=0, jbnd=21, kbnd=0), dict()),
]


def make(storage="f32", members=MEMBERS):
    doms = []
    for m, (dims, bnd, ex) in enumerate(members):
        d = Domain(19, dims, storage, **PHYS, **bnd)
        f0 = inputs.perturbed_uniform_flow(19, dims, seed=50 + m, order=3)
        d.set_state(f0)
        if bnd["ibnd"] == 1:
            d.set_inflow(0.04, ex.get("udir", 0.0))
            d.set_inflow_series(1, np.linspace(0.04, 0.05, 80), np.full(80, ex.get("udir", 0.0)))
        if "box" in ex:
            d.set_blanking(inputs.box_mask(dims, *ex["box"]))
        if ex.get("turb"):
            tu = inputs.turbulence_fields(60 + m, dims[1], dims[2], 7)
            d.set_turbulence(tu[0], tu[1], tu[2], 7, 1e-3)
        if ex.get("force"):
            nx, ny, nz = dims
            F = np.zeros((nz + 2, ny + 2, nx + 2, 3), np.float32)
            F[..., 0] = 1e-6
            F[..., 2] = -5e-7
            d.set_external_forcing(F)
        doms.append(d)
    return doms


def states(doms):
    return [d.get_state() for d in doms]


def assert_same(a, b, what):
    for m, (x, y) in enumerate(zip(a, b)):
        assert x.shape == y.shape
        if not np.array_equal(x, y):
            d = np.abs(x.astype(np.float64) - y.astype(np.float64))
            pytest.fail(f"{what}: member {m} differs, max |df| {d.max():.3e} at {np.unravel_index(d.argmax(), d.shape)}")


@pytest.fixture(scope="module")
def reference_runs():
    """Each member stepped alone with lbm_step: states after 1, 9 and 45 steps."""
    doms = make()
    out = {}
    done = 0
    for n in (1, 9, 45):
        for d in doms:
            d.step(1 + done, n - done)
        done = n
        out[n] = states(doms)
    for d in doms:
        d.close()
    return out


@pytest.mark.parametrize("launch", [(256, 0), (64, 0), (1024, 2), (128, 1)])
def test_ensemble_matches_single_domain_steps(reference_runs, launch):
    print(f"ensemble of {len(MEMBERS)} members, launch {launch}", flush=True)
    doms = make()
    ens = Ensemble(doms)
    ens.set_launch(*launch)
    ens.step(0 + 1, 1)          # the in-place first step runs per member
    assert_same(states(doms), reference_runs[1], "after 1 step")
    ens.step(2, 8)              # one graph of 8 steps
    assert_same(states(doms), reference_runs[9], "after 9 steps")
    ens.step(10, 20)            # graphs of 16 + 4 steps
    ens.step(30, 16)            # the 16-step graph replayed again
    assert_same(states(doms), reference_runs[45], "after 45 steps")
    ens.close()
    for d in doms:
        d.close()


def test_ensemble_plane_arrangement(reference_runs, monkeypatch):
    """Members in the arrangement of population buffers of 4 GiB and more (one plane per
    population), stepped by the batched kernels: the default-arrangement single-domain states."""
    monkeypatch.setenv("LBM_LAYOUT", "planes")
    monkeypatch.setenv("LBM_STEP_VARIANT", "3")
    doms = make()
    ens = Ensemble(doms)
    ens.step(1, 1)
    ens.step(2, 8)
    assert_same(states(doms), reference_runs[9], "after 9 steps (planes)")
    ens.close()
    for d in doms:
        d.close()


def test_tuner_restores_state_and_keeps_results(reference_runs):
    doms = make()
    ens = Ensemble(doms)
    ens.step(1, 9)
    before = states(doms)
    block, bpc, ms = ens.tune(10, trial_steps=3)
    print(f"tuned launch: block {block}, blocks/CU {bpc}, {ms * 1e3:.1f} us/step", flush=True)
    assert block in (64, 128, 256, 512, 1024) and bpc in (0, 1, 2, 4, 8) and ms > 0
    assert ens.launch() == (block, bpc)
    assert_same(states(doms), before, "after tuning")
    ens.step(10, 36)
    assert_same(states(doms), reference_runs[45], "tuned launch, 45 steps")
    ens.close()


def test_members_stay_usable_between_ensemble_steps():
    """A member's inputs changed between ensemble steps are picked up (tables are rebuilt)."""
    a, b = make(), make()
    ea = Ensemble(a)
    ea.step(1, 5)
    for d in b:
        d.step(1, 5)
    for d in (a[0], b[0]):
        d.set_inflow(0.03)
    ea.step(6, 5)
    for d in b:
        d.step(6, 5)
    assert_same(states(a), states(b), "after an input change")
    ea.close()


def test_fp16_storage_ensemble():
    members = MEMBERS[:3]
    a, b = make("f16s", members), make("f16s", members)
    ens = Ensemble(a)
    ens.step(1, 25)
    for d in b:
        d.step(1, 25)
    assert_same(states(a), states(b), "f16s")
    ens.close()


def test_invalid_members_are_rejected():
    d1 = Domain(19, (16, 8, 8), **PHYS, kbnd=11)
    d2 = Domain(19, (16, 8, 8), ibgk=2, kinevisc=1e-3)
    d3 = Domain(19, (16, 8, 8), **PHYS, jbnd=11, kbnd=11)
    d4 = Domain(27, (16, 8, 8), **PHYS)
    for bad, msg in (([d1, d2], "same lattice"), ([d1, d4], "same lattice"), ([d3], "at most one closed axis"),
                     ([d1, d1], "repeats")):
        with pytest.raises(LBMError, match=msg):
            Ensemble(bad)
    with pytest.raises(LBMError, match="no state"):
        Ensemble([d1]).step(0, 1)


def test_more_steps_than_one_table_of_inflow_values():
    """One call of more than 1024 steps: the per-step inflow values are uploaded in tables of 1024
    steps, the buffer parity runs on across them.  Small inflow members (inflow planes folded into
    the batched step), with a time series of the inflow speed that changes every step."""
    members = [((12, 8, 6), dict(ibnd=1, jbnd=0, kbnd=12), dict(udir=10.0)),
               ((10, 6, 8), dict(ibnd=1, jbnd=0, kbnd=0), dict()),
               ((8, 8, 6), dict(ibnd=0, jbnd=0, kbnd=11), dict())]
    nsteps = 1100
    series = np.linspace(0.03, 0.05, nsteps + 2)
    print(f"ensemble of {len(members)} members, {nsteps} steps in one call", flush=True)

    def build():
        doms = make(members=members)
        for d, (_, bnd, ex) in zip(doms, members):
            if bnd["ibnd"] == 1:
                d.set_inflow_series(1, series, np.full(nsteps + 2, ex.get("udir", 0.0)))
        return doms

    a = build()
    for d in a:
        d.step(1, nsteps)
    ref = states(a)
    for d in a:
        d.close()
    b = build()
    ens = Ensemble(b)
    ens.step(1, nsteps)
    ens.synchronize()
    ens.close()
    assert_same(states(b), ref, "1100 steps in one call")
    for d in b:
        d.close()
