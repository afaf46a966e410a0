"""Domains tiled along j (jbnd = LBM_BND_HALO) with the halo exchange done through
lbm_halo_pack / lbm_halo_unpack, the way one process per GPU would run them with RCCL send/recv
between the two calls (src/m_mpi_halo_exchange_j.F90:55-120).  On one device the exchange is a
device buffer handed from one context to the other.  A periodic-j domain split into tiles must
give the untiled run's state bit for bit: the interior, the ghost planes of i and k, and the
crossing populations of the j halos; so must split steps (lbm_step_part: the cells next to the
halos, the pack, then the other cells while the transfer runs on another stream).
"""
import numpy as np
import pytest
import torch

from lbm_amd import inputs
from lbm_amd.domain import BND_HALO, PART_EDGE, PART_INNER, Domain
from lbm_amd.lattice import lattice

pytestmark = pytest.mark.gpu

Q = 27


def cy_of(q):
    return lattice(q).cy


def tile_of(arr, j0, j1, ny, axis):
    """Planes j0-1 .. j1+1 of a record with periodic j (ghosts taken from the wrapped interior)."""
    idx = [(j - 1) % ny + 1 for j in range(j0 - 1, j1 + 2)]
    return np.take(arr, idx, axis=axis)


def make(dims, storage, codes, o, halo, f0, mask, tu, j0=1, j1=None, q=Q):
    nx, ny, nz = dims
    j1 = ny if j1 is None else j1
    sub = (nx, j1 - j0 + 1, nz)
    dom = Domain(q, sub, storage, ibnd=codes[0], jbnd=BND_HALO if halo else codes[1], kbnd=codes[2], **o)
    dom.set_state(tile_of(f0, j0, j1, ny, 1) if halo else f0)
    dom.set_inflow(0.04, 10.0)
    if mask is not None:
        dom.set_blanking(tile_of(mask, j0, j1, ny, 1) if halo else mask)
    if tu is not None and codes[0] == 1:
        dom.set_turbulence(*(np.ascontiguousarray(t[:, :, j0 - 1:j1]) for t in tu[:3]), 10, 1e-3)
    return dom


def exchange(tiles, bufs):
    """Periodic neighbours: tile t's lower neighbour is t-1, its upper t+1 (mod the count)."""
    n = len(tiles)
    for t, d in enumerate(tiles):
        d.halo_pack(0, bufs[t][0].data_ptr())
        d.halo_pack(1, bufs[t][1].data_ptr())
    for d in tiles:
        d.synchronize()
    for t, d in enumerate(tiles):
        d.halo_unpack(0, bufs[(t - 1) % n][1].data_ptr())   # ghost j = 0 <- plane ny of the tile below
        d.halo_unpack(1, bufs[(t + 1) % n][0].data_ptr())   # ghost j = ny+1 <- plane 1 of the tile above
    for d in tiles:
        d.synchronize()


def exchange_both(tiles, bufs):
    """The same with one pack and one unpack launch per tile (side 2): bufs[t] = (send, receive),
    each two halos long, side 0's first."""
    n = len(tiles)
    for t, d in enumerate(tiles):
        d.halo_pack(2, bufs[t][0].data_ptr())
    for d in tiles:
        d.synchronize()
    nb = bufs[0][0].numel() // 2
    for t in range(n):
        bufs[t][1][:nb].copy_(bufs[(t - 1) % n][0][nb:])   # ghost j = 0 <- plane ny of the tile below
        bufs[t][1][nb:].copy_(bufs[(t + 1) % n][0][:nb])   # ghost j = ny+1 <- plane 1 of the tile above
    torch.cuda.synchronize()
    for t, d in enumerate(tiles):
        d.halo_unpack(2, bufs[t][1].data_ptr())
    for d in tiles:
        d.synchronize()


def split_step(tiles, bufs, it, streams):
    """Step it (what step(it - 1, 1) runs) as lbm_step_part's two launches with the transfer during
    the inner one: per tile the
    edge rows and the pack (both sides) on the tile's stream, an event; the inner rows; the
    transfer on a stream of its own after every tile's event; each tile's unpack on its stream
    after the transfer."""
    n = len(tiles)
    nb = bufs[0][0].numel() // 2
    packed = []
    for t, d in enumerate(tiles):
        d.step_part(it, PART_EDGE)
        d.halo_pack(2, bufs[t][0].data_ptr())
        ev = torch.cuda.Event()
        ev.record(streams[t])
        packed.append(ev)
        d.step_part(it, PART_INNER)
    side = torch.cuda.current_stream()
    for ev in packed:
        side.wait_event(ev)
    for t in range(n):
        bufs[t][1][:nb].copy_(bufs[(t - 1) % n][0][nb:])   # ghost j = 0 <- plane ny of the tile below
        bufs[t][1][nb:].copy_(bufs[(t + 1) % n][0][:nb])   # ghost j = ny+1 <- plane 1 of the tile above
    moved = torch.cuda.Event()
    moved.record(side)
    for t, d in enumerate(tiles):
        streams[t].wait_event(moved)
        d.halo_unpack(2, bufs[t][1].data_ptr())


CASES = {
    "inflow_closedk_solids_turb": ((1, 0, 12), "f32", dict(ibgk=3, ihrr=1, ivreman=1, smagorinsky=0.16733)),
    "inflow_closedk_f16s": ((1, 0, 21), "f16s", dict(ibgk=3, ihrr=1, ivreman=1, smagorinsky=0.16733)),
    "periodic_i_closedk_bgk2": ((0, 0, 22), "f32", dict(ibgk=2, ihrr=0, ivreman=0)),
    "inflow_periodic_k": ((1, 0, 0), "f32", dict(ibgk=3, ihrr=1, ivreman=0)),
    "closed_i_periodic_k": ((11, 0, 0), "f64", dict(ibgk=3, ihrr=1, ivreman=1, smagorinsky=0.16733)),
    "all_periodic": ((0, 0, 0), "f32", dict(ibgk=3, ihrr=1, ivreman=1, smagorinsky=0.16733)),
    "q19_inflow_closedk": ((1, 0, 12), "f32", dict(ibgk=3, ihrr=1, ivreman=1, smagorinsky=0.16733, q=19)),
}


def run_case(name, splits, both=False, split=False):
    codes, storage, o = CASES[name]
    o = dict(o, kinevisc=2e-3)
    q = o.pop("q", Q)
    dims, steps = (18, 13, 10), 20
    nx, ny, nz = dims
    # a start whose j ghost planes are the periodic copies, as the reference's records hold them
    f0 = tile_of(inputs.perturbed_uniform_flow(q, dims, seed=21, u0=(0.04, 0.012, -0.01), order=3), 1, ny, ny, 1)
    # solids next to the inflow plane and the lower wall, across the tile boundary j = 5|6
    mask = inputs.box_mask(dims, (1, 4, 1), (3, 7, 2))
    tu = inputs.turbulence_fields(21, ny, nz, 10)
    print(f"halo case {name} storage={storage} tiles={splits}", flush=True)
    ref = make(dims, storage, codes, o, False, f0, mask, tu, q=q)
    ref.step(1, steps)
    fr = ref.get_state()
    ref.close()
    tiles = [make(dims, storage, codes, o, True, f0, mask, tu, j0, j1, q=q) for j0, j1 in splits]
    nb = tiles[0].halo_bytes() * (2 if both or split else 1)
    bufs = [[torch.empty(nb, dtype=torch.uint8, device="cuda") for _ in range(2)] for _ in tiles]
    streams = [torch.cuda.ExternalStream(d.stream_ptr()) for d in tiles]
    for it in range(1, steps + 1):
        if split:
            split_step(tiles, bufs, it + 1, streams)   # the step that d.step(it, 1) runs
            continue
        for d in tiles:
            d.step(it, 1)
        (exchange_both if both else exchange)(tiles, bufs)
    for d in tiles:
        d.synchronize()
    states = [d.get_state() for d in tiles]
    for d in tiles:
        d.close()
    for (j0, j1), f in zip(splits, states):
        assert np.array_equal(f[:, 1:-1], fr[:, j0:j1 + 1]), (name, j0, j1)
    if len(splits) == 1:
        # the record's ghost planes (the previous step's boundary pass): for a tile that is its own
        # neighbour, the crossing populations of its halos are the periodic run's ghost values
        cys = cy_of(q)
        f = states[0]
        assert np.array_equal(f[:, 0][..., cys == 1], fr[:, 0][..., cys == 1])
        assert np.array_equal(f[:, -1][..., cys == -1], fr[:, -1][..., cys == -1])


@pytest.mark.parametrize("name", list(CASES))
def test_self_exchange_matches_periodic(name):
    """One tile whose neighbours above and below are itself: a periodic domain."""
    run_case(name, [(1, 13)])


@pytest.mark.parametrize("name", list(CASES))
def test_two_tiles_match_untiled(name):
    """Two tiles of unequal height (j = 1..5 and 6..13) exchanging every step."""
    run_case(name, [(1, 5), (6, 13)])


@pytest.mark.parametrize("name", ["inflow_closedk_solids_turb", "closed_i_periodic_k"])
def test_two_tiles_both_sides_per_launch(name):
    run_case(name, [(1, 6), (7, 13)], both=True)


@pytest.mark.parametrize("name", ["inflow_closedk_solids_turb", "inflow_periodic_k"])
def test_two_tiles_separate_inflow_pass(name, monkeypatch):
    """The inflow/outflow planes written by their own kernel after the step (LBM_INFLOW_FOLD=0):
    the face's copies from inflow ghosts are then made by the step kernel itself."""
    monkeypatch.setenv("LBM_INFLOW_FOLD", "0")
    run_case(name, [(1, 5), (6, 13)])


@pytest.mark.parametrize("splits", [[(1, 4), (5, 9), (10, 13)], [(1, 8), (9, 13)]], ids=["4-5-4", "8-5"])
@pytest.mark.parametrize("name", list(CASES))
def test_split_steps_match_untiled(name, splits):
    """Edge rows, pack, inner rows with the transfer on another stream, unpack: the untiled run bit
    for bit.  Two edge rows per side with a closed face, one without.  Tiles of 4 rows (with a
    face: one launch, the edge part is the whole tile), 5 and 8."""
    run_case(name, splits, split=True)


@pytest.mark.parametrize("name", ["inflow_closedk_solids_turb", "closed_i_periodic_k"])
def test_two_tiles_plane_arrangement(name, monkeypatch):
    """Halo pack/unpack and the tiles' steps in the arrangement of population buffers of 4 GiB and
    more (one plane per population, 64-bit plane offsets), forced on a small grid."""
    monkeypatch.setenv("LBM_LAYOUT", "planes")
    monkeypatch.setenv("LBM_STEP_VARIANT", "3")
    run_case(name, [(1, 5), (6, 13)])


def test_three_tiles_fp16():
    run_case("inflow_closedk_f16s", [(1, 4), (5, 9), (10, 13)])


def test_halo_errors():
    d = Domain(Q, (12, 6, 6), "f32", ibnd=1, jbnd=0, kbnd=12)
    with pytest.raises(RuntimeError, match="not a halo domain"):
        d.halo_bytes()
    d.close()
    with pytest.raises(RuntimeError, match="at most one closed axis"):
        Domain(Q, (12, 6, 6), "f32", ibnd=11, jbnd=BND_HALO, kbnd=12)
    d = Domain(Q, (12, 6, 6), "f32", ibnd=1, jbnd=BND_HALO, kbnd=12)
    assert d.halo_bytes() == 9 * 14 * 8 * 4
    buf = torch.empty(d.halo_bytes(), dtype=torch.uint8, device="cuda")
    with pytest.raises(RuntimeError, match="no state"):
        d.halo_pack(0, buf.data_ptr())
    with pytest.raises(RuntimeError, match="side must be"):
        d.set_state(inputs.perturbed_uniform_flow(Q, (12, 6, 6), seed=1, u0=(0.02, 0, 0), order=3))
        d.halo_unpack(3, buf.data_ptr())
    d.close()


def test_split_step_errors():
    f = inputs.perturbed_uniform_flow(Q, (12, 6, 6), seed=3, u0=(0.02, 0, 0), order=3)
    d = Domain(Q, (12, 6, 6), "f32", ibnd=1, jbnd=0, kbnd=12)
    d.set_state(f)
    with pytest.raises(RuntimeError, match="halo domain"):
        d.step_part(1, PART_EDGE)
    d.close()
    d = Domain(Q, (12, 6, 6), "f32", ibnd=1, jbnd=BND_HALO, kbnd=12)
    d.set_state(f)
    with pytest.raises(RuntimeError, match="no split step is open"):
        d.step_part(1, PART_INNER)
    with pytest.raises(RuntimeError, match="part must be"):
        d.step_part(1, 3)
    d.step_part(1, PART_EDGE)
    mask = np.zeros((8, 8, 14), dtype=np.uint8)   # (nz+2, ny+2, nx+2)
    for call in (lambda: d.step(1, 1), lambda: d.step_part(1, PART_EDGE), d.get_state,
                 lambda: d.set_state(f), lambda: d.set_blanking(mask), lambda: d.set_inflow(0.02, 0.0),
                 d.totals, d.macrovars):
        with pytest.raises(RuntimeError, match="split step is open"):
            call()
    with pytest.raises(RuntimeError, match="differs"):
        d.step_part(2, PART_INNER)
    d.step_part(1, PART_INNER)
    d.step(1, 1)
    assert np.isfinite(d.get_state()).all()
    d.close()
