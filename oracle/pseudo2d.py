"""Oracle (test infrastructure only): numpy restatement of the inflow-turbulence generator that
lbm_generate_turbulence runs on the device.  It follows the reference's
inflow_turbulence_compute (src/m_inflow_turbulence_compute.F90:3-62) and pseudo2D
(src/m_pseudo2D.F90:3-167): analytic spectrum with r1 = r2 = sqrt(8)/rx (r2 formed from rx as
in src/m_pseudo2D.F90:85), c = sqrt(1/(dk*summ)) with the mean mode excluded
(src/m_pseudo2D.F90:99-104), fampl(1,0,0) = fampl(2,0,0) = 0, the FFTW c2r transform (here
numpy's irfft2 times n1*n2: the same unnormalized inverse with the imaginary parts of the l = 0
and l = n1/2 columns discarded), then AR(1) with 0.98 and per-slice normalization of slices
1..nrturb.  The random phases use the same SplitMix64 counters as the device code (the
reference draws them with random_number, so the phases themselves are not comparable with it:
parity of the phases is unpinned; everything downstream of them is pinned by
tests/test_oracle_turbulence.py).
"""
from __future__ import annotations

import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from lbm_amd.inputs import uniform01  # noqa: E402

PI = 3.141592653589   # the reference's constant (src/m_pseudo2D.F90:44)


def spectrum_constants(n1, n2, rx=10.0 / np.sqrt(3.0)):
    pi2 = 2.0 * PI
    dk = pi2 ** 2 / (n1 * n2)
    kappa, lam = pi2 / n1, pi2 / n2
    r1 = np.sqrt(8.0) / rx
    r2 = np.sqrt(8.0) / rx
    l = np.arange(-n1 // 2 + 1, n1 // 2 + 1)
    p = np.arange(-n2 // 2 + 1, n2 // 2 + 1)
    summ = np.exp(-2.0 * (kappa ** 2 * l[None, :] ** 2 / r1 ** 2 + lam ** 2 * p[:, None] ** 2 / r2 ** 2)).sum() - 1.0
    c = np.sqrt(1.0 / (dk * summ))
    return dk, c, kappa ** 2 / r1 ** 2, lam ** 2 / r2 ** 2


def raw_fields(ny, nz, nfields, seed):
    """pseudo2D fields before the AR(1) step: (nfields, nz, ny); field index = variable*ns + slice."""
    n1, n2 = ny + (ny & 1), nz + (nz & 1)
    nh = n1 // 2 + 1
    dk, c, ak, al = spectrum_constants(n1, n2)
    phi = 2.0 * np.pi * uniform01(seed, nfields * nh * n2).reshape(nfields, nh, n2)
    l = np.arange(nh)[:, None]
    pp = np.arange(n2)[None, :]
    p = np.where(pp < n2 // 2, pp, pp - n2)
    e = np.exp(-(ak * l ** 2 + al * p ** 2))
    X = e * np.sqrt(dk) * c * np.exp(1j * phi)          # (nf, nh, n2)
    X[:, 0, 0] = 0.0
    X = np.transpose(X, (0, 2, 1))                      # (nf, n2, nh): p rows, l halved last
    y = np.fft.irfft2(X, s=(n2, n1), axes=(-2, -1)) * (n1 * n2)
    return y[:, :nz, :ny]


def ar1_normalize(f, timecor=0.98):
    """src/m_inflow_turbulence_compute.F90:36-60 on f of shape (nrturb+1, nz, ny), in place."""
    n = f.shape[1] * f.shape[2]
    b = np.sqrt(1.0 - timecor ** 2)
    for i in range(1, f.shape[0]):
        f[i] = timecor * f[i - 1] + b * f[i]
        f[i] -= f[i].sum() / n
        f[i] /= np.sqrt((f[i] ** 2).sum() / (n - 1))
    return f


def generate(ny, nz, nrturb, seed):
    """(uu, vv, ww, rr), each (nrturb+1, nz, ny), FP64."""
    ns = nrturb + 1
    raw = raw_fields(ny, nz, 4 * ns, seed).reshape(4, ns, nz, ny)
    return tuple(ar1_normalize(raw[v].copy()) for v in range(4))
