"""Workloads built on the C-ABI: the batched-ensemble members (D3Q19 regularized + Vreman,
inflow/outflow i, periodic j, no-slip k = 1, free-slip k = nz, a box obstacle and seeded inflow
forcing per member)."""
from __future__ import annotations

import numpy as np

from . import inputs
from .domain import Domain

ENSEMBLE_PHYSICS = dict(ibgk=3, ihrr=1, ivreman=1, smagorinsky=0.16733, kinevisc=1e-4, ibnd=1, jbnd=0, kbnd=12)


def ensemble_domains(specs, storage="f32", nrturb=100, ampl=1e-3, uini=0.05, it0=0, physics=None):
    """One Domain per spec of inputs.ensemble_specs, started from the equilibrium of perturbed
    moments (the reference's cold start, src/main.F90:243-247)."""
    phys = dict(ENSEMBLE_PHYSICS, **(physics or {}))
    doms = []
    for sp in specs:
        dims = sp["dims"]
        d = Domain(19, dims, storage, **phys)
        rho, u, v, w = inputs.perturbed_moments(dims, sp["seed"], u0=(uini, 0.0, 0.0))
        d.set_inflow(uini)
        d.set_blanking(inputs.box_mask(dims, sp["box_lo"], sp["box_hi"]))
        tu = inputs.turbulence_fields(sp["seed"], dims[1], dims[2], nrturb)
        d.set_turbulence(tu[0], tu[1], tu[2], nrturb, ampl)
        d.init_equilibrium(rho, u, v, w, it0)
        doms.append(d)
    return doms


def cells(specs):
    return int(sum(np.prod(sp["dims"]) for sp in specs))
