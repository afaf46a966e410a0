"""Long-run statistics of the cylinder-wake physics (D3Q27 regularized + Vreman, inflow/outflow i,
periodic j, no-slip k=1 / free-slip k=nz, the reference's cylinder r=7 at (50, ny/2)) at
96x32x16 over 2000 steps, run in restart segments of 50 steps, sampled at every segment end of
the second half.  Shared by the fixture writer (the reference's own FP32 and FP64 builds) and the
GPU test (the HIP library, restarted at the same steps through lbm_set_state).
"""
from __future__ import annotations

import numpy as np

Q = 27
DIMS = (96, 32, 16)
STEPS = 2000
SEGMENT = 50
SEED = 31
OPTS = dict(ibgk=3, ihrr=1, ivreman=1, smagorinsky=0.16733, ibnd=1, jbnd=0, kbnd=12, uini=0.05, udir=0.0,
            visc=1e-4)
WAKE_I = 70   # plane behind the cylinder for the wake profile


def initial_state(dtype=np.float32):
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
    from lbm_amd import inputs
    return inputs.perturbed_uniform_flow(Q, DIMS, seed=SEED, u0=(OPTS["uini"], 0.0, 0.0), order=3, dtype=dtype)


def blanking():
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
    from lbm_amd import inputs
    return inputs.cylinder_mask(DIMS, 50, DIMS[1] // 2, 7)


def velocities(f):
    """u, v, w of the interior cells, (nz, ny, nx), from a restart record (nz+2, ny+2, nx+2, Q)."""
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
    from lbm_amd.lattice import lattice
    lat = lattice(Q)
    g = f[1:-1, 1:-1, 1:-1].astype(np.float64)
    rho = g.sum(-1)
    return (g * lat.cx).sum(-1) / rho, (g * lat.cy).sum(-1) / rho, (g * lat.cz).sum(-1) / rho


def statistics(samples, solid):
    """samples: list of (u, v, w); solid: (nz, ny, nx) bool.  Returns the mean streamwise velocity
    per k over the fluid cells, the turbulent kinetic energy per k (time variance at each fluid cell,
    averaged over i and j) and the time-mean u across j at i = WAKE_I, averaged over k."""
    u = np.stack([s[0] for s in samples])
    v = np.stack([s[1] for s in samples])
    w = np.stack([s[2] for s in samples])
    fluid = ~solid
    um = u.mean(0)
    tke = 0.5 * (u.var(0) + v.var(0) + w.var(0))
    n = fluid.sum(axis=(1, 2))
    u_k = (um * fluid).sum(axis=(1, 2)) / n
    tke_k = (tke * fluid).sum(axis=(1, 2)) / n
    u_wake = um[:, :, WAKE_I - 1].mean(0)
    return dict(u_k=u_k, tke_k=tke_k, u_wake_j=u_wake)


def sample_steps():
    return [n for n in range(SEGMENT, STEPS + 1, SEGMENT) if n > STEPS // 2]
