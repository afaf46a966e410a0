"""Seeded synthetic inputs for the lattice update.

The reference seeds its initial state from the compiler's ``random_number`` with seeds taken
from the clock (src/m_set_random_seed3.F90:36, src/m_inipert.F90:121-130), which is not
reproducible across compilers, and synthesises its inflow-turbulence fields with FFTW
(src/m_pseudo2D.F90).  Inputs here are therefore produced by a project-written generator
(SplitMix64 -> U[0,1), computed in FP64 and rounded to the storage type) and handed to both
the reference and this library through the reference's restart-file layout, so that both
consume identical bytes.

Array conventions (C-contiguous numpy):
  populations f  : (nz+2, ny+2, nx+2, nl)  == reference f(nl,0:nx+1,0:ny+1,0:nz+1)
  cell fields    : (nz+2, ny+2, nx+2)      == reference rho(0:nx+1,0:ny+1,0:nz+1)
  turbulence     : (nrturb+1, nz, ny)      == reference uu(ny,nz,0:nrturb)
"""
from __future__ import annotations

import numpy as np

from .lattice import lattice

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """The n outputs of the SplitMix64 stream started at ``seed`` (skipping ``offset``)."""
    with np.errstate(over="ignore"):
        idx = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """U[0,1) in FP64 from the top 53 bits of SplitMix64."""
    return (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Standard normals by Box-Muller on consecutive uniform pairs."""
    u = uniform01(seed, 2 * n, 2 * offset)
    u1 = 1.0 - u[0::2]
    u2 = u[1::2]
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def feq_fp64(q: int, rho: np.ndarray, u: np.ndarray, v: np.ndarray, w: np.ndarray, order: int) -> np.ndarray:
    """Hermite equilibrium in FP64 (second order, or third order with the full H3 term as the
    reference evaluates it, src/m_postcoll_kernel.F90:136-183).  Returns (..., nl)."""
    lat = lattice(q)
    cu = (lat.cx[None, :] * u[..., None] + lat.cy[None, :] * v[..., None] + lat.cz[None, :] * w[..., None])
    usq = (u * u + v * v + w * w)[..., None]
    r = rho[..., None]
    feq = r * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq)
    if order == 3:
        rest = (lat.cx == 0) & (lat.cy == 0) & (lat.cz == 0)
        third = 4.5 * r * cu * (cu * cu - usq)
        feq = feq + np.where(rest[None, :], 0.0, third)
    return lat.w[None, :] * feq if feq.ndim == 2 else lat.w * feq


def fill_ghosts(f: np.ndarray) -> np.ndarray:
    """Ghost = clamped interior cell (src/m_fillghosts.F90:191-197).  In place."""
    f[0] = f[1]
    f[-1] = f[-2]
    f[:, 0] = f[:, 1]
    f[:, -1] = f[:, -2]
    f[:, :, 0] = f[:, :, 1]
    f[:, :, -1] = f[:, :, -2]
    return f


def perturbed_uniform_flow(q: int, dims, seed: int, u0=(0.05, 0.0, 0.0), amp: float = 1e-3,
                           order: int = 2, dtype=np.float32, rho0: float = 1.0) -> np.ndarray:
    """rho = rho0, u = u0 + U(-amp, amp) per component per cell, f = feq(rho, u) (FP64, then
    rounded), ghosts by clamped copy.  Shape (nz+2, ny+2, nx+2, q)."""
    nx, ny, nz = dims
    n = nx * ny * nz
    r = uniform01(seed, 3 * n).reshape(3, nz, ny, nx)
    vel = [u0[c] + amp * (2.0 * r[c] - 1.0) for c in range(3)]
    rho = np.full((nz, ny, nx), rho0)
    fi = feq_fp64(q, rho.reshape(-1), vel[0].reshape(-1), vel[1].reshape(-1), vel[2].reshape(-1), order)
    f = np.zeros((nz + 2, ny + 2, nx + 2, q), dtype=np.float64)
    f[1:-1, 1:-1, 1:-1] = fi.reshape(nz, ny, nx, q)
    fill_ghosts(f)
    return f.astype(dtype)


def smooth_field(seed: int, ny: int, nz: int, corr: float, offset: int = 0) -> np.ndarray:
    """Unit-variance Gaussian-correlated field on (nz, ny), periodic, correlation length
    ``corr`` cells: white noise filtered with exp(-(k*corr)^2/4) in Fourier space."""
    noise = normal(seed, ny * nz, offset).reshape(nz, ny)
    ky = np.fft.fftfreq(ny) * 2.0 * np.pi
    kz = np.fft.fftfreq(nz) * 2.0 * np.pi
    filt = np.exp(-((kz[:, None] * corr) ** 2 + (ky[None, :] * corr) ** 2) / 4.0)
    g = np.real(np.fft.ifft2(np.fft.fft2(noise) * filt))
    g = g - g.mean()
    s = g.std()
    return g / s if s > 1e-12 else np.zeros_like(g)


def turbulence_fields(seed: int, ny: int, nz: int, nrturb: int, corr: float = 8.0,
                      ar1: float = 0.98, dtype=np.float32):
    """uu, vv, ww, rr of shape (nrturb+1, nz, ny): smooth unit-variance fields with AR(1)
    time correlation ``ar1`` between consecutive slices (the reference's generator uses the
    same structure, src/m_inflow_turbulence_compute.F90:35-60)."""
    out = []
    for c in range(4):
        fld = np.empty((nrturb + 1, nz, ny))
        prev = smooth_field(seed * 4 + c, ny, nz, corr, offset=0)
        fld[0] = prev
        b = np.sqrt(1.0 - ar1 * ar1)
        for t in range(1, nrturb + 1):
            prev = ar1 * prev + b * smooth_field(seed * 4 + c, ny, nz, corr, offset=t * ny * nz)
            fld[t] = prev
        out.append(fld.astype(dtype))
    return tuple(out)


def cylinder_mask(dims, ic: int, jc: int, radius: int) -> np.ndarray:
    """Solid cylinder along k: blanked where (i-ic)**2 + (j-jc)**2 < radius**2 for interior
    i, j and every k including ghosts (the reference's test, src/m_cylinder.F90:13-17).
    Returns uint8 (nz+2, ny+2, nx+2)."""
    nx, ny, nz = dims
    i = np.arange(nx + 2)[None, :]
    j = np.arange(ny + 2)[:, None]
    inside = ((i - ic) ** 2 + (j - jc) ** 2 < radius ** 2)
    inside[:, 0] = inside[:, -1] = False
    inside[0, :] = inside[-1, :] = False
    return np.broadcast_to(inside[None], (nz + 2, ny + 2, nx + 2)).astype(np.uint8).copy()


def box_mask(dims, lo, hi) -> np.ndarray:
    """Solid box [lo, hi] (inclusive, 1-based interior indices i,j,k)."""
    nx, ny, nz = dims
    m = np.zeros((nz + 2, ny + 2, nx + 2), np.uint8)
    m[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = 1
    return m


def ensemble_specs(seed: int, n: int, sizes_i=(96, 128, 160), sizes_j=(64, 96), sizes_k=(64, 96),
                   box_side=(8, 16)):
    """Seeded members of the batched-ensemble workload: dims drawn uniformly from the size sets,
    a solid box of side s in [box_side] standing on the k = 1 wall at i0 in [nx/6, nx/3],
    j0 anywhere, and a per-member seed for the initial perturbation and the inflow forcing."""
    u = uniform01(seed, 8 * n).reshape(n, 8)
    out = []
    for m in range(n):
        r = u[m]
        nx = sizes_i[int(r[0] * len(sizes_i))]
        ny = sizes_j[int(r[1] * len(sizes_j))]
        nz = sizes_k[int(r[2] * len(sizes_k))]
        s = box_side[0] + int(r[3] * (box_side[1] - box_side[0] + 1))
        i0 = nx // 6 + int(r[4] * (nx // 3 - nx // 6))
        j0 = 1 + int(r[5] * (ny - s))
        out.append(dict(dims=(nx, ny, nz), box_lo=(i0, j0, 1), box_hi=(i0 + s - 1, j0 + s - 1, s),
                        seed=seed * 1000 + m + 1))
    return out


def perturbed_moments(dims, seed: int, u0=(0.05, 0.0, 0.0), amp: float = 1e-3, dtype=np.float32):
    """rho = 1 and u = u0 + U(-amp, amp) per component on every cell including ghosts, for
    lbm_init_equilibrium.  Shapes (nz+2, ny+2, nx+2)."""
    nx, ny, nz = dims
    n = (nx + 2) * (ny + 2) * (nz + 2)
    r = uniform01(seed, 3 * n).reshape(3, nz + 2, ny + 2, nx + 2)
    rho = np.ones((nz + 2, ny + 2, nx + 2), dtype)
    vel = [(u0[c] + amp * (2.0 * r[c] - 1.0)).astype(dtype) for c in range(3)]
    return rho, vel[0], vel[1], vel[2]
