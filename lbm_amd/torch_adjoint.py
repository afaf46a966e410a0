"""PyTorch harness of the hand-written adjoint (lbm_adjoint_run).

``LatticeKE`` is a torch.autograd.Function whose forward runs the lattice update on the GPU
through the C ABI and returns the time-mean kinetic energy J over a box, and whose backward
returns the adjoint's dJ/d(body-force field) and dJ/d(inflow shear profile).  ``ForcingProblem``
builds those inputs from parameters with ordinary torch operations, so torch's autograd chains
the adjoint's field gradients to the parameters:

    body force  F_x(i,j,k) = sum_b a_b exp(-|x - x_b|^2 / (2 sigma^2))   (j periodic)
    inflow      uvel_shear(k) = sum_p c_p zeta_k^(p/7),  zeta_k = (k - 1/2)/nz

Each evaluation starts from the same state: the seeded initial populations, one step with the
baseline inputs (zero force, the 1/7 power law), then nsteps steps with the given inputs.
"""
from __future__ import annotations

import numpy as np
import torch

from . import inputs
from .domain import Domain


class LatticeKE(torch.autograd.Function):
    @staticmethod
    def forward(ctx, force, shear, problem):
        J, gF, gS = problem.evaluate(force.detach().cpu().numpy(), shear.detach().cpu().numpy(), adjoint=True)
        ctx.save_for_backward(torch.from_numpy(gF).to(force.device), torch.from_numpy(gS).to(shear.device))
        return force.new_tensor(J)

    @staticmethod
    def backward(ctx, gout):
        gF, gS = ctx.saved_tensors
        return gout * gF, gout * gS, None


class ForcingProblem:
    """D3Q19 regularized + Vreman, inflow/outflow i, periodic j, no-slip k = 1 (free-slip k = nz),
    Gaussian body-force blobs near i = i_blobs and a power-law inflow profile; objective: the
    time-mean kinetic energy over i in box_i (all j, k)."""

    def __init__(self, dims=(192, 96, 96), storage="f32", nsteps=400, every=20, box_i=(140, 180), nblobs=64,
                 sigma=4.0, i_blobs=40, ncoef=8, seed=5, kinevisc=1e-3, uini=0.05, smagorinsky=0.16733,
                 bnd=(1, 0, 12)):
        self.dims, self.storage, self.nsteps, self.every = tuple(dims), storage, nsteps, every
        nx, ny, nz = self.dims
        self.box = (tuple(box_i), (1, ny), (1, nz))
        self.uini = uini
        self.dom = Domain(19, self.dims, storage, ibgk=3, ihrr=1, ivreman=1, smagorinsky=smagorinsky,
                          kinevisc=kinevisc, ibnd=bnd[0], jbnd=bnd[1], kbnd=bnd[2])
        self.f0 = inputs.perturbed_uniform_flow(19, self.dims, seed=seed, u0=(uini, 0.0, 0.0), order=3,
                                                dtype=np.float32 if storage == "f32" else np.float64)
        r = inputs.uniform01(seed + 77, 3 * nblobs).reshape(3, nblobs)
        self.centres = np.stack([i_blobs + 8.0 * (2 * r[0] - 1), 1 + (ny - 1) * r[1], 8 + (nz - 16) * r[2]], 1)
        self.sigma = sigma
        self.ncoef = ncoef
        zeta = (np.arange(1, nz + 1) - 0.5) / nz
        self.basis = torch.from_numpy(np.stack([zeta ** (p / 7.0) for p in range(ncoef)]))   # (ncoef, nz)
        self.c0 = torch.zeros(ncoef, dtype=torch.float64)
        self.c0[1] = 1.0
        self.a0 = torch.from_numpy(2e-5 * (2 * inputs.uniform01(seed + 91, nblobs) - 1))
        self.evaluations = 0

    # -- parameters -> fields (torch, differentiable) ------------------------------------------
    def force_field(self, a):
        """(nz+2, ny+2, nx+2, 3) float64 tensor; x-component of the blobs in the interior."""
        nx, ny, nz = self.dims
        i = torch.arange(1, nx + 1, dtype=torch.float64)
        j = torch.arange(1, ny + 1, dtype=torch.float64)
        k = torch.arange(1, nz + 1, dtype=torch.float64)
        c = torch.from_numpy(self.centres)
        s2 = 2.0 * self.sigma ** 2
        dj = j[None, :] - c[:, 1:2]
        dj = dj - ny * torch.round(dj / ny)                 # periodic in j
        gx = torch.exp(-(i[None, :] - c[:, 0:1]) ** 2 / s2)  # (nb, nx)
        gy = torch.exp(-dj ** 2 / s2)                        # (nb, ny)
        gz = torch.exp(-(k[None, :] - c[:, 2:3]) ** 2 / s2)  # (nb, nz)
        fx = torch.einsum("b,bk,bj,bi->kji", a, gz, gy, gx)
        F = torch.zeros((nz + 2, ny + 2, nx + 2, 3), dtype=torch.float64)
        F = F.index_put((torch.arange(1, nz + 1)[:, None, None], torch.arange(1, ny + 1)[None, :, None],
                         torch.arange(1, nx + 1)[None, None, :], torch.tensor(0)), fx)
        return F

    def shear_profile(self, coef):
        return coef @ self.basis

    # -- one evaluation -----------------------------------------------------------------------
    def evaluate(self, force, shear, adjoint=True):
        d = self.dom
        d.set_external_forcing(None)
        d.set_inflow(self.uini, 0.0, self.shear_profile(self.c0).numpy())
        d.set_state(self.f0)
        d.step(0, 1)
        d.set_external_forcing(force)
        d.set_inflow(self.uini, 0.0, shear)
        self.evaluations += 1
        if adjoint:
            return d.adjoint(1, self.nsteps, self.every, self.box)
        return d.objective(1, self.nsteps, self.box)

    def directional(self, a, coef, da, dc):
        """Tangent run (storage "f64t"): (J, dJ) with dJ the derivative of J along (da, dc),
        carried through the forward kernels in forward-mode FP64."""
        with torch.no_grad():
            F = self.force_field(a).numpy()
            dF = self.force_field(da).numpy()
            s = self.shear_profile(coef).numpy()
            ds = self.shear_profile(dc).numpy()
        d = self.dom
        d.set_external_forcing(None)
        d.set_inflow(self.uini, 0.0, self.shear_profile(self.c0).numpy())
        d.set_state(self.f0)
        d.step(0, 1)
        d.set_external_forcing(F)
        d.set_inflow(self.uini, 0.0, s)
        d.set_tangent(dF, ds)
        J = d.objective(1, self.nsteps, self.box)
        return J, d.objective_tangent()

    def J(self, a, coef):
        """Objective as a differentiable function of the blob amplitudes and profile coefficients."""
        return LatticeKE.apply(self.force_field(a), self.shear_profile(coef), self)

    def J_value(self, a, coef):
        with torch.no_grad():
            return self.evaluate(self.force_field(a).numpy(), self.shear_profile(coef).numpy(), adjoint=False)

    def close(self):
        self.dom.close()
