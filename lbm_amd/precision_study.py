#!/usr/bin/env python3
"""Storage-precision study: the same run with populations stored as FP32, weight-shifted FP16
and weight-shifted BF16 (FP32 arithmetic), compared with FP64 storage (FP64 arithmetic) after
given step counts.

    python -m lbm_amd.precision_study [--physics bgk|regvre] [--dims 768 256 192] [--steps 100 1000]

Physics "bgk": D3Q19 BGK, tau = 0.6, periodic i and j, no-slip k, u = (0.05,0,0) plus a seeded
uniform perturbation of 1e-3.  Physics "regvre": D3Q27 third-order regularized + Vreman
(c = 0.07), nu = 1e-4, inflow/outflow i, periodic j, no-slip k=1 / free-slip k=nz, a cylinder
of radius dims[1]/16 on the k axis at (dims[0]/4, dims[1]/2), seeded inflow forcing at i = 10.
Every run starts from the equilibrium of the same FP64 moments (lbm_init_equilibrium).

Per field (rho - 1, u, v, w, and the velocity 3-vector): relative L2 error
||q - q64|| / ||q64|| over interior cells.  Mean profile: the i,j-mean of u per k; its deviation
is max_k |<u>(k) - <u64>(k)| / max_k |<u64>(k)|.  TKE profile: 0.5 <u'^2 + v'^2 + w'^2>(k) with
primes about the i,j-means; deviation as for the mean profile.  Prints one JSON object.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import inputs
from .domain import Domain


def physics(name, dims):
    nx, ny, nz = dims
    if name == "bgk":
        return 19, dict(ibgk=2, ihrr=0, ivreman=0, kinevisc=(0.6 - 0.5) / 3.0, ibnd=0, jbnd=0, kbnd=11), None
    r = max(4, ny // 16)
    return 27, dict(ibgk=3, ihrr=1, ivreman=1, smagorinsky=math.sqrt(0.07 / 2.5), kinevisc=1e-4, ibnd=1, jbnd=0,
                    kbnd=12), (nx // 4, ny // 2, r)


def run(storage, q, opts, dims, cyl, moments, turb, checkpoints, nrturb):
    d = Domain(q, dims, storage, **opts)
    if opts["ibnd"] == 1:
        d.set_inflow(0.05)
    if cyl is not None:
        d.set_blanking(inputs.cylinder_mask(dims, *cyl))
    if turb is not None:
        d.set_turbulence(turb[0], turb[1], turb[2], nrturb, 1e-3)
    d.init_equilibrium(*moments, it0=0)
    out = {}
    it = 0
    t0 = time.perf_counter()
    for n in checkpoints:
        d.step(it, n - it)
        it = n
        r, u, v, w = (a[1:-1, 1:-1, 1:-1].astype(np.float64) for a in d.macrovars(np.float64))
        out[n] = (r, u, v, w)
        print(f"[precision] {storage} step {n} ({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    d.close()
    return out


def rel_l2(a, b):
    den = math.sqrt(float((b ** 2).sum()))
    return math.sqrt(float(((a - b) ** 2).sum())) / den if den > 0 else float("nan")


def profiles(u, v, w):
    # arrays (nz, ny, nx): means over j, i per k
    mu, mv, mw = (a.mean(axis=(1, 2)) for a in (u, v, w))
    tke = 0.5 * (((u - mu[:, None, None]) ** 2) + ((v - mv[:, None, None]) ** 2) + ((w - mw[:, None, None]) ** 2)).mean(axis=(1, 2))
    return mu, tke


def compare(x, ref):
    r, u, v, w = x
    r0, u0, v0, w0 = ref
    vel = math.sqrt(float(((u - u0) ** 2 + (v - v0) ** 2 + (w - w0) ** 2).sum())) / math.sqrt(float((u0 ** 2 + v0 ** 2 + w0 ** 2).sum()))
    mu, tke = profiles(u, v, w)
    mu0, tke0 = profiles(u0, v0, w0)
    return {"rho-1": rel_l2(r - 1.0, r0 - 1.0), "u": rel_l2(u, u0), "v": rel_l2(v, v0), "w": rel_l2(w, w0),
            "velocity": vel, "mean_profile": float(np.abs(mu - mu0).max() / np.abs(mu0).max()),
            "tke_profile": float(np.abs(tke - tke0).max() / max(np.abs(tke0).max(), 1e-300))}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--physics", default="regvre", choices=["bgk", "regvre"])
    ap.add_argument("--dims", type=int, nargs=3, default=(768, 256, 192))
    ap.add_argument("--steps", type=int, nargs="+", default=(100, 1000))
    ap.add_argument("--storages", nargs="+", default=("f32", "f16s", "bf16s"))
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--nrturb", type=int, default=250)
    a = ap.parse_args(argv)
    dims = tuple(a.dims)
    q, opts, cyl = physics(a.physics, dims)
    print(f"[precision] {a.physics} q={q} {dims}: building inputs", file=sys.stderr, flush=True)
    moments = inputs.perturbed_moments(dims, a.seed, dtype=np.float64)
    turb = None
    if opts["ibnd"] == 1:
        turb = inputs.turbulence_fields(a.seed, dims[1], dims[2], a.nrturb, dtype=np.float64)
    cps = sorted(a.steps)
    ref = run("f64", q, opts, dims, cyl, moments, turb, cps, a.nrturb)
    res = {"physics": a.physics, "q": q, "dims": list(dims), "steps": cps, "errors": {}}
    for st in a.storages:
        x = run(st, q, opts, dims, cyl, moments, turb, cps, a.nrturb)
        res["errors"][st] = {str(n): compare(x[n], ref[n]) for n in cps}
        res.setdefault("finite", {})[st] = bool(all(np.isfinite(x[n][1]).all() for n in cps))
        del x
    print(json.dumps(res))
    return res


if __name__ == "__main__":
    main()
