"""The oracle (oracle/lbm_oracle.c) pinned by what the mathematics fixes, independently of the
golden vectors: equilibrium moments, the regularization projection and the Malaspinas
recursion, Vreman special cases, the forcing impulse and its sign, streaming as a permutation,
conservation, and the viscosity of BGK (Taylor-Green decay).  All in the FP64 build."""
import ctypes as C

import numpy as np
import pytest

import oracle
from lbm_amd.lattice import lattice

pytestmark = []


def one_cell(q, **o):
    """A 1x1x1 periodic FP64 domain; returns (sim, helpers)."""
    sim = oracle.OracleSim(q, (2, 2, 2), "f64", **o)
    return sim


def fields(sim):
    return sim.rho, sim.u, sim.v, sim.w


def H2(lat):
    c = lat.c.astype(float)
    return np.einsum("lp,lq->lpq", c, c) - np.eye(3)[None] / 3.0


def H3(lat):
    c = lat.c.astype(float)
    d = np.eye(3)
    t = np.einsum("lp,lq,lr->lpqr", c, c, c)
    t -= (np.einsum("lp,qr->lpqr", c, d) + np.einsum("lq,pr->lpqr", c, d) + np.einsum("lr,pq->lpqr", c, d)) / 3.0
    return t


def feq_of(sim, rho, u):
    """Run or_fequil3 on a 2x2x2 domain with uniform moments; return the feq of one cell."""
    sim.rho[...] = rho
    sim.u[...], sim.v[...], sim.w[...] = u
    f = np.zeros_like(sim.fA)
    sim.lib.or_fequil3(C.byref(sim.p), oracle._ptr(f), oracle._ptr(sim.rho), oracle._ptr(sim.u),
                       oracle._ptr(sim.v), oracle._ptr(sim.w))
    return f[1, 1, 1].copy()


@pytest.mark.parametrize("q", [19, 27])
def test_lattice_isotropy(q):
    lat = lattice(q)
    c = lat.c.astype(float)
    w = lat.w
    assert np.isclose(w.sum(), 1.0)
    assert np.allclose(w @ c, 0)
    assert np.allclose(np.einsum("l,lp,lq->pq", w, c, c), np.eye(3) / 3)
    m4 = np.einsum("l,lp,lq,lr,ls->pqrs", w, c, c, c, c)
    d = np.eye(3)
    iso = (np.einsum("pq,rs->pqrs", d, d) + np.einsum("pr,qs->pqrs", d, d) + np.einsum("ps,qr->pqrs", d, d)) / 9
    assert np.allclose(m4, iso)
    # opposite directions are the adjacent pairs (2,3),(4,5),... (1-based)
    for l in range(1, q, 2):
        assert np.array_equal(c[l], -c[l + 1])
    # Hermite orthogonality
    assert np.allclose(np.einsum("l,lpq->pq", w, H2(lat)), 0)
    assert np.allclose(np.einsum("l,lr,lpq->rpq", w, c, H2(lat)), 0)
    assert np.allclose(np.einsum("l,lpqr->pqr", w, H3(lat)), 0)


@pytest.mark.parametrize("q,ibgk", [(19, 2), (19, 3), (27, 2), (27, 3)])
def test_equilibrium_moments(q, ibgk):
    """sum feq = rho, sum c feq = rho u, sum c c feq = rho u u + rho I/3 (any order), and the
    third-order term changes only odd third moments."""
    lat = lattice(q)
    sim = one_cell(q, ibgk=ibgk)
    rho, u = 1.013, np.array([0.041, -0.017, 0.008])
    f = feq_of(sim, rho, u)
    c = lat.c.astype(float)
    assert np.isclose(f.sum(), rho, rtol=0, atol=1e-14)
    assert np.allclose(f @ c, rho * u, rtol=0, atol=1e-14)
    assert np.allclose(np.einsum("l,lp,lq->pq", f, c, c), rho * np.outer(u, u) + rho * np.eye(3) / 3, atol=1e-14)
    f2 = feq_of(one_cell(q, ibgk=2), rho, u)
    d3 = f - f2
    if ibgk == 3:
        # 4.5 w rho cu (cu^2 - u.u), zero at rest (the reference's H3:A3 contraction)
        cu = c @ u
        expect = 4.5 * lat.w * rho * cu * (cu ** 2 - u @ u)
        expect[0] = 0.0
        assert np.allclose(d3, expect, atol=1e-15)
    else:
        assert np.array_equal(d3, np.zeros(q))


def postcoll_one(q, f_cell, rho, u, **o):
    sim = one_cell(q, **o)
    sim.fA[...] = f_cell
    sim.rho[...] = rho
    sim.u[...], sim.v[...], sim.w[...] = u
    sim.tau[...] = 3.0 * sim.kinevisc + 0.5
    abl = np.zeros(2)
    sim.lib.or_postcoll(C.byref(sim.p), oracle._ptr(sim.fA), oracle._ptr(sim.tau), oracle._ptr(sim.rho),
                        oracle._ptr(sim.u), oracle._ptr(sim.v), oracle._ptr(sim.w), oracle._ptr(abl))
    return sim.fA[1, 1, 1].copy(), sim.tau[1, 1, 1]


@pytest.mark.parametrize("q", [19, 27])
def test_regularization_projection_and_recursion(q):
    """R(fneq): its H2 moment equals the input's (projection), and for D3Q27 its xyz H3 moment
    is the Malaspinas value u_x a2_yz + u_y a2_xz + u_z a2_xy."""
    lat = lattice(q)
    rng = np.random.default_rng(0)
    rho, u = 1.02, np.array([0.03, -0.02, 0.01])
    tau = 0.8
    nu = (tau - 0.5) / 3.0
    feq = feq_of(one_cell(q, ibgk=3), rho, u)
    fneq = 1e-3 * rng.standard_normal(q) * lat.w
    fout, _ = postcoll_one(q, feq + fneq, rho, u, ibgk=3, ihrr=1, kinevisc=nu)
    fac = 1.0 - 1.0 / tau
    rneq = (fout - feq) / fac
    a2_in = np.einsum("lpq,l->pq", H2(lat), fneq)
    a2_out = np.einsum("lpq,l->pq", H2(lat), rneq)
    assert np.allclose(a2_out, a2_in, atol=1e-15)
    # idempotence: regularizing R(fneq) again returns it
    fout2, _ = postcoll_one(q, feq + rneq, rho, u, ibgk=3, ihrr=1, kinevisc=nu)
    assert np.allclose((fout2 - feq) / fac, rneq, atol=1e-15)
    if q == 27:
        m_xyz = np.einsum("l,l->", H3(lat)[:, 0, 1, 2], rneq)
        expect = u[0] * a2_in[1, 2] + u[1] * a2_in[0, 2] + u[2] * a2_in[0, 1]
        assert np.isclose(m_xyz, expect, rtol=1e-12, atol=1e-18)


def vreman_tau(q, alpha, nu=1e-3, smag=0.2):
    """tau the oracle forms for a cell whose non-equilibrium tensor is alpha (fneq built so that
    sum H2 fneq = alpha, non-regularized path)."""
    lat = lattice(q)
    rho, u = 1.0, np.zeros(3)
    feq = feq_of(one_cell(q, ibgk=2), rho, u)
    fneq = 4.5 * lat.w * np.einsum("lpq,pq->l", H2(lat), alpha)
    _, tau = postcoll_one(q, feq + fneq, rho, u, ibgk=2, ihrr=0, ivreman=1, smagorinsky=smag, kinevisc=nu)
    return tau


def test_vreman_special_cases():
    nu, smag = 1e-3, 0.2
    cst = 2.5 * smag * smag
    a = 1e-4
    shear = np.array([[0, a, 0], [a, 0, 0], [0, 0, 0.0]])
    # beta = alpha^T alpha: b11 = b22 = a^2 -> B = a^4, |alpha|^2 = 2 a^2 -> nu_t = c a / sqrt(2)
    assert np.isclose(vreman_tau(27, shear, nu, smag), 3 * (nu + cst * a / np.sqrt(2)) + 0.5, rtol=1e-9)
    # no strain: nu_t = 0
    assert np.isclose(vreman_tau(27, np.zeros((3, 3)), nu, smag), 3 * nu + 0.5, rtol=1e-14)
    # rank-one alpha (one-dimensional strain): B = 0 up to eps -> nu_t = 0
    assert np.isclose(vreman_tau(19, np.diag([a, 0.0, 0.0]), nu, smag), 3 * nu + 0.5, rtol=1e-12)
    # a non-symmetric-looking index error would break this: alpha with two shear components
    b = np.array([[0, a, 2 * a], [a, 0, 0], [2 * a, 0, 0.0]])
    bt = b.T @ b
    B = bt[0, 0] * bt[1, 1] - bt[0, 1] ** 2 + bt[0, 0] * bt[2, 2] - bt[0, 2] ** 2 + bt[1, 1] * bt[2, 2] - bt[1, 2] ** 2
    nut = cst * np.sqrt(B / (b ** 2).sum())
    assert np.isclose(vreman_tau(27, b, nu, smag), 3 * (nu + nut) + 0.5, rtol=1e-9)


@pytest.mark.parametrize("q", [19, 27])
def test_forcing_impulse_and_sign(q):
    """forcings_apply adds feq2(rho, u - F/rho) - feq2(rho, u): no mass, momentum -F
    (the sign as written in src/m_forcings_apply_kernel.F90:188-190)."""
    lat = lattice(q)
    sim = oracle.OracleSim(q, (3, 2, 2), "f64", ibgk=2)
    rng = np.random.default_rng(1)
    sim.fA[...] = lat.w * (1 + 0.01 * rng.standard_normal(sim.fA.shape))
    sim.rho[...] = 1.0 + 0.01 * rng.random(sim.rho.shape)
    sim.u[...] = 0.03 * rng.standard_normal(sim.u.shape)
    F = 1e-4 * rng.standard_normal(sim.rho.shape + (3,))
    before = sim.fA.copy()
    sim.lib.or_forcings_apply(C.byref(sim.p), oracle._ptr(sim.fA), oracle._ptr(F), oracle._ptr(sim.rho),
                              oracle._ptr(sim.u), oracle._ptr(sim.v), oracle._ptr(sim.w))
    df = (sim.fA - before)[1:-1, 1:-1, 1:-1]
    assert np.allclose(df.sum(-1), 0, atol=1e-15)
    mom = np.einsum("...l,lp->...p", df, lat.c.astype(float))
    assert np.allclose(mom, -F[1:-1, 1:-1, 1:-1], rtol=1e-9, atol=1e-18)


def test_drift_is_the_pull_permutation():
    q, dims = 27, (5, 4, 3)
    lat = lattice(q)
    sim = oracle.OracleSim(q, dims, "f64")
    rng = np.random.default_rng(2)
    fin = rng.random(sim.fA.shape)
    fout = np.zeros_like(fin)
    sim.lib.or_drift(C.byref(sim.p), oracle._ptr(fout), oracle._ptr(fin))
    nx, ny, nz = dims
    for l in range(q):
        cx, cy, cz = lat.c[l]
        src = fin[1 - cz:nz + 1 - cz, 1 - cy:ny + 1 - cy, 1 - cx:nx + 1 - cx, l]
        assert np.array_equal(fout[1:-1, 1:-1, 1:-1, l], src)


@pytest.mark.parametrize("q,o", [(19, dict(ibgk=2)), (27, dict(ibgk=3, ihrr=1, ivreman=1, smagorinsky=0.17))])
def test_periodic_box_conserves_mass_and_momentum(q, o):
    from lbm_amd import inputs
    dims = (8, 6, 5)
    f0 = inputs.perturbed_uniform_flow(q, dims, seed=3, u0=(0.04, -0.01, 0.02), amp=0.01, dtype=np.float64)
    sim = oracle.OracleSim(q, dims, "f64", kinevisc=0.01, **o)
    sim.restart(f0)
    lat = lattice(q)

    def totals(f):
        a = f[1:-1, 1:-1, 1:-1]
        return a.sum(), (a.reshape(-1, q).sum(0)) @ lat.c.astype(float)

    m0, p0 = totals(f0)
    f = sim.run(1, 31)
    m1, p1 = totals(f)
    assert abs(m1 - m0) <= 1e-11 * m0
    assert np.allclose(p1, p0, atol=1e-12)


def test_taylor_green_viscosity():
    """2-D Taylor-Green vortex (periodic, D3Q19 BGK): kinetic energy decays as
    exp(-4 nu k^2 t), nu = (tau - 1/2)/3."""
    q, n = 19, 32
    dims = (n, n, 2)
    lat = lattice(q)
    from lbm_amd import inputs
    x = np.arange(n) + 0.5
    X, Y = np.meshgrid(x, x, indexing="xy")
    k = 2 * np.pi / n
    U = 0.01
    u = U * np.sin(k * X) * np.cos(k * Y)
    v = -U * np.cos(k * X) * np.sin(k * Y)
    nz = 2
    uu = np.broadcast_to(u, (nz, n, n)).reshape(-1)
    vv = np.broadcast_to(v, (nz, n, n)).reshape(-1)
    rho = np.ones_like(uu)
    fi = inputs.feq_fp64(q, rho, uu, vv, np.zeros_like(uu), 2).reshape(nz, n, n, q)
    f0 = np.zeros((nz + 2, n + 2, n + 2, q))
    f0[1:-1, 1:-1, 1:-1] = fi
    inputs.fill_ghosts(f0)
    tau = 0.8
    nu = (tau - 0.5) / 3
    sim = oracle.OracleSim(q, dims, "f64", ibgk=2, kinevisc=nu)
    sim.restart(f0)

    def energy():
        _, uo, vo, _ = (a[1:-1, 1:-1, 1:-1] for a in sim.macros())
        return (uo ** 2 + vo ** 2).mean()

    sim.run(1, 101)          # past the start-up transient of an equilibrium initial state
    e0 = energy()
    steps = 200
    sim.run(101, 101 + steps)
    e1 = energy()
    expect = np.exp(-4 * nu * k * k * steps)
    assert abs(e1 / e0 / expect - 1) < 1e-2   # O(k^2) lattice dispersion; a wrong nu(tau) is off by far more


def test_fillghosts_is_the_clamped_copy():
    """or_fillghosts (src/m_fillghosts.F90:19-26) on a ragged tiny grid of distinct values: every
    ghost holds its clamped interior cell, i.e. numpy's edge padding of the interior, and the
    interior is untouched."""
    for q, dims in ((19, (3, 2, 4)), (27, (2, 5, 3))):
        sim = oracle.OracleSim(q, dims, "f64")
        rng = np.random.default_rng(q)
        f = rng.random(sim.fA.shape)
        inner = f[1:-1, 1:-1, 1:-1].copy()
        sim.lib.or_fillghosts(C.byref(sim.p), oracle._ptr(f))
        expect = np.pad(inner, ((1, 1), (1, 1), (1, 1), (0, 0)), mode="edge")
        assert np.array_equal(f, expect)


def feq2_np(lat, rho, u):
    """Second-order equilibrium w rho (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u.u), FP64, per cell."""
    cu = np.einsum("...p,lp->...l", u, lat.c.astype(float))
    usq = (u * u).sum(-1)[..., None]
    return lat.w * rho[..., None] * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usq)


def feq3_np(lat, rho, u):
    """feq2 plus the third-order term 4.5 w rho c.u ((c.u)^2 - u.u) (src/m_postcoll_kernel.F90)."""
    cu = np.einsum("...p,lp->...l", u, lat.c.astype(float))
    usq = (u * u).sum(-1)[..., None]
    return feq2_np(lat, rho, u) + 4.5 * lat.w * rho[..., None] * cu * (cu * cu - usq)


def forcing_moment_errors(df, rho, u, F, lat):
    """Deviations of the increment df from what feq2(rho, v) - feq2(rho, u) with v = u - F/rho
    fixes (src/m_forcings_apply_kernel.F90:142-207): second moments rho (v v - u u), the third
    moments m_xxy = rho (v_y - u_y)/3 = -F_y/3 (any third-order term adds rho (v_x^2 v_y - u_x^2 u_y))
    and, on D3Q27, m_xyz = 0 (a third-order term adds rho (v_x v_y v_z - u_x u_y u_z))."""
    c = lat.c.astype(float)
    v = u - F / rho[..., None]
    pi = np.einsum("...l,lp,lq->...pq", df, c, c)
    pi_expect = rho[..., None, None] * (np.einsum("...p,...q->...pq", v, v) - np.einsum("...p,...q->...pq", u, u))
    m_xxy = np.einsum("...l,l->...", df, c[:, 0] ** 2 * c[:, 1])
    m_xyz = np.einsum("...l,l->...", df, c[:, 0] * c[:, 1] * c[:, 2])
    return (np.abs(pi - pi_expect).max() / np.abs(pi_expect).max(),
            np.abs(m_xxy + F[..., 1] / 3).max() / np.abs(F[..., 1] / 3).max(),
            np.abs(m_xyz).max() / np.abs(F).max())


@pytest.mark.parametrize("q", [19, 27])
def test_forcing_uses_the_second_order_equilibrium(q):
    """The moments of the forcing increment pin the second-order equilibrium (also with ibgk = 3,
    README.md:81): a slip to the third-order form, planted here in numpy, fails the same check."""
    lat = lattice(q)
    sim = oracle.OracleSim(q, (3, 2, 2), "f64", ibgk=3)
    rng = np.random.default_rng(5)
    sim.fA[...] = lat.w * (1 + 0.01 * rng.standard_normal(sim.fA.shape))
    sim.rho[...] = 1.0 + 0.05 * rng.random(sim.rho.shape)
    sim.u[...], sim.v[...], sim.w[...] = (0.05 * rng.standard_normal(sim.u.shape) for _ in range(3))
    F = 1e-4 * rng.standard_normal(sim.rho.shape + (3,))
    before = sim.fA.copy()
    sim.lib.or_forcings_apply(C.byref(sim.p), oracle._ptr(sim.fA), oracle._ptr(F), oracle._ptr(sim.rho),
                              oracle._ptr(sim.u), oracle._ptr(sim.v), oracle._ptr(sim.w))
    inner = (slice(1, -1),) * 3
    df = (sim.fA - before)[inner]
    rho = sim.rho[inner]
    u = np.stack([sim.u[inner], sim.v[inner], sim.w[inner]], -1)
    Fi = F[inner]
    e_pi, e_xxy, e_xyz = forcing_moment_errors(df, rho, u, Fi, lat)
    assert e_pi < 1e-9 and e_xxy < 1e-8 and e_xyz < 1e-12
    # the planted slip: feq3 in place of feq2
    v = u - Fi / rho[..., None]
    e3 = forcing_moment_errors(feq3_np(lat, rho, v) - feq3_np(lat, rho, u), rho, u, Fi, lat)
    assert e3[1] > 1e-4 or e3[2] > 1e-4
    # and the restated increment itself agrees with numpy's feq2 difference
    assert np.allclose(df, feq2_np(lat, rho, v) - feq2_np(lat, rho, u), rtol=0, atol=1e-15)
