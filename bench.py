#!/usr/bin/env python3
"""Benchmark of the fused lattice update on one MI355X.

Workload (the configuration the headline metric is quoted on): D3Q27, third-order Hermite
equilibrium, regularized collision + Vreman SGS (c = 0.07, i.e. smagorinsky = sqrt(0.07/2.5)),
FP32, 384x192x128 cells, inflow u=0.05 at i=1 / outflow at i=nx, periodic j, no-slip k=1,
free-slip k=nz, solid cylinder of radius 12 along k at (96,96), smooth seeded inflow forcing at
slice i=10 (amplitude 1e-3, correlation length 8), nu = 1e-4, 1000 steps by default.

A "step" is one pass of the whole time-loop body (fused stream-collide kernel + boundary
kernels) over the grid.  Inputs are synthetic and resident in HBM before timing starts.

    python bench.py [--steps K] [--warmup W] [--storage f32|f16s|bf16s|f64] [--no-cpu-baseline]

Prints ONE JSON line.  --gpus N > 1 is refused: this build targets one GPU.

    python bench.py --workload ensemble [--members 64] [--steps 500] [--small-members]

times the batched ensemble instead: 64 D3Q19 regularized + Vreman members with sizes drawn by
seed from {96,128,160}x{64,96}x{64,96}, a seeded box obstacle (side 8-16) and seeded inflow
forcing each, FP32, graph-captured batched steps with the autotuned launch shape, against the
same members stepped with one launch sequence per member; --small-members draws the sizes from
{24,32,40}x{16,24}x{16,24} (box side 4-8) instead.

    python bench.py --workload adjoint [--steps 400]

times the hand-written adjoint (forward with checkpoints every 20 steps, recomputation, reverse
sweep) of D3Q19 regularized + Vreman, FP32, 192x96x96, inflow/outflow i, periodic j, no-slip
k=1, with 64 Gaussian body-force blobs and an 8-coefficient power-law inflow profile, objective
the time-mean kinetic energy over i in [140,180]; the gradient is checked against FP64 central
differences (h = 1e-6) on a subset of the parameters and against the FP64 adjoint on all.

    python bench.py --workload tile [--tiles 8] [--split]

times one GPU's j-tile of the headline workload (tile 4 of 8, 384x24x128) with its halo exchange
every step; --split steps it in two launches with the transfer during the second.

    python bench.py --workload q19 [--storage f32|f16s]

times D3Q19 BGK (tau = 0.6, periodic i and j, no-slip k) at 768x256x192.

    python bench.py --workload large [--storage f32|f16s|bf16s]

times the headline physics at the storage-study size 768x256x192 (D3Q27 FP32 populations: 5.2 GB
per buffer, row-blocked and addressed per slab of the slowest axis).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SMAG = math.sqrt(0.07 / 2.5)


def workload(storage, dims=(384, 192, 128)):
    return dict(q=27, dims=tuple(dims), storage=storage, ibgk=3, ihrr=1, ivreman=1, smagorinsky=SMAG,
                kinevisc=1e-4, ibnd=1, jbnd=0, kbnd=12, uini=0.05, udir=0.0, cyl=(96, 96, 12), ip=10,
                ampl=1e-3, corr=8.0, seed=2)


def bytes_per_cell(q, storage):
    s = {"f32": 4, "f64": 8, "f16s": 2, "bf16s": 2}[storage]
    return 2 * q * s


def make_inputs(w, nrturb):
    from lbm_amd import inputs
    dt = np.float64 if w["storage"] == "f64" else np.float32
    f0 = inputs.perturbed_uniform_flow(w["q"], w["dims"], seed=w["seed"], u0=(w["uini"], 0, 0), order=3, dtype=dt)
    nx, ny, nz = w["dims"]
    turb = inputs.turbulence_fields(w["seed"], ny, nz, nrturb, corr=w["corr"], dtype=dt)
    ic, jc, r = w["cyl"]
    mask = inputs.cylinder_mask(w["dims"], ic, jc, r)
    return f0, turb, mask


def traffic_from_profiles(name):
    """HBM bytes per launch measured with rocprofv3 PMC counters (profiles/<name>.json)."""
    p = os.path.join(ROOT, "profiles", name + ".json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as fh:
            return json.load(fh).get("traffic_bytes_per_launch")
    except Exception:
        return None


def default_cpu_threads():
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    return min(n, int(env)) if env.isdigit() and int(env) > 0 else n


def cpu_baseline(w, f0, turb, sample_steps, threads, sample_steps_1t=0):
    """The reference's own OpenMP build (oracle/_ref, amdflang) timed on this host on a bounded
    sample of the same workload, by its "Total TS routines" timer (src/m_wtime.F90:75).  The build
    carries the workload's cylinder (r=12 at (96,96)) in place of the reference's hard-coded one;
    with sample_steps_1t > 0 a second, shorter run on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import build_ref  # noqa: E402
    import oracle  # noqa: E402
    ic, jc, r = w["cyl"]
    exe = build_ref.exe_path(w["q"], "sp", w["dims"], omp=True, cylinder=(ic, jc, r))
    if not os.path.exists(exe):
        try:
            exe = build_ref.build(w["q"], "sp", w["dims"], omp=True, cylinder=(ic, jc, r))
        except Exception:
            return None
    nx, ny, nz = w["dims"]

    def run(steps, nthreads):
        nr = steps + 5
        tf = tuple(x[: nr + 1] for x in turb)
        wd = tempfile.mkdtemp(prefix="lbm_cpu_")
        try:
            _, out = oracle.run_reference(exe, f0, nt0=1, nt1=1 + steps, turb_fields=tf, workdir=wd,
                                          read_output=False, threads=nthreads, ibgk=3, ihrr=1, ivreman=1,
                                          smagorinsky=float(np.float32(SMAG)), ibnd=1, jbnd=0, kbnd=12,
                                          uini=w["uini"], udir=0.0, visc=w["kinevisc"], ampl=w["ampl"], nrturb=nr,
                                          experiment="cylinder", ltiming=True, ltesting=False, ldump=False)
        finally:
            shutil.rmtree(wd, ignore_errors=True)
        m = re.search(r"Total TS routines\s+time\s*=\s*([0-9.Ee+-]+)", out)
        return float(m.group(1)) if m else None

    secs = run(sample_steps, threads)
    if secs is None:
        return None
    res = {"value": round(nx * ny * nz * sample_steps / secs / 1e6, 3), "unit": "MLUPS", "cores": threads,
           "kind": "reference",
           "sample": f"{sample_steps} steps of the same 384x192x128 D3Q27 reg+Vreman workload (same inputs, "
                     f"cylinder r={r} at ({ic},{jc})), reference built with amdflang -O2 -fopenmp, "
                     f"OMP_NUM_THREADS={threads}, 'Total TS routines' timer {secs:.2f} s",
           "affinity_cpus": len(os.sched_getaffinity(0))}
    if sample_steps_1t > 0:
        s1 = run(sample_steps_1t, 1)
        if s1:
            res["one_thread"] = {"value": round(nx * ny * nz * sample_steps_1t / s1 / 1e6, 3), "unit": "MLUPS",
                                 "cores": 1, "sample": f"{sample_steps_1t} steps, OMP_NUM_THREADS=1, {s1:.2f} s"}
    return res


def bench_ensemble(a):
    from lbm_amd import inputs, workloads
    from lbm_amd.domain import Ensemble
    # --small-members: the same physics on members of 24-40 x 16-24 x 16-24 cells, where one launch
    # per member is launch-bound rather than bandwidth-bound
    size_sets = (dict(sizes_i=(24, 32, 40), sizes_j=(16, 24), sizes_k=(16, 24), box_side=(4, 8)) if a.small_members
                 else dict(sizes_i=(96, 128, 160), sizes_j=(64, 96), sizes_k=(64, 96), box_side=(8, 16)))
    specs = inputs.ensemble_specs(4, a.members, **size_sets)
    cells = workloads.cells(specs)
    print(f"[bench] building {a.members} ensemble members, {cells / 1e6:.1f} M cells", file=sys.stderr, flush=True)
    doms = workloads.ensemble_domains(specs, a.storage, nrturb=100)
    ens = Ensemble(doms)
    it = 0
    ens.step(it, a.warmup)
    it += a.warmup
    ens.synchronize()
    print("[bench] autotuning the launch shape", file=sys.stderr, flush=True)
    block, bpc, tune_ms = ens.tune(it, trial_steps=5)
    print(f"[bench] tuned: block {block}, blocks/CU {bpc}, {tune_ms:.3f} ms/step", file=sys.stderr, flush=True)
    # one launch sequence per member and step (the members' own contexts)
    n_pm = max(10, a.steps // 5)
    for d in doms:
        d.synchronize()
    t0 = time.perf_counter()
    for s_ in range(n_pm):
        for d in doms:
            d.step(it + s_, 1)
    for d in doms:
        d.synchronize()
    t_pm = time.perf_counter() - t0
    it += n_pm
    # batched
    print(f"[bench] timed {a.steps} batched steps", file=sys.stderr, flush=True)
    ens.set_timing(True)
    ens.synchronize()
    t0 = time.perf_counter()
    ens.step(it, a.steps)
    ens.synchronize()
    secs = time.perf_counter() - t0
    ev_ms, ev_steps = ens.timing_read()
    finite = all(bool(np.isfinite(d.totals()[0])) for d in doms)
    ens.close()
    for d in doms:
        d.close()
    mlups = cells * a.steps / secs / 1e6
    mlups_pm = cells * n_pm / t_pm / 1e6
    step_ms = ev_ms / max(ev_steps, 1)
    bpc_bytes = bytes_per_cell(19, a.storage)
    achieved = bpc_bytes * cells / (step_ms * 1e-3) / 1e9
    line = {
        "metric": "MLUPS (batched ensemble, D3Q19 reg+Vreman)", "value": round(mlups, 1), "unit": "MLUPS",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(secs / a.steps * 1e3, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"f32": "f32", "f64": "f64", "f16s": "f16-store/f32", "bf16s": "bf16-store/f32"}[a.storage],
        "data": "synthetic (seeded sizes, box obstacles, perturbed moments and inflow forcing)",
        "config": {"workload": f"{a.members} D3Q19 regularized + Vreman members, sizes from "
                               + "x".join("{" + ",".join(str(v) for v in size_sets[k]) + "}"
                                          for k in ("sizes_i", "sizes_j", "sizes_k"))
                               + f", box obstacle side {size_sets['box_side'][0]}-{size_sets['box_side'][1]}, "
                               "inflow/outflow i, periodic j, no-slip k=1 / free-slip k=nz, inflow forcing at i=10, "
                               "nu=1e-4",
                   "cells": cells, "members": a.members, "storage": a.storage, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     # measured for the default ensemble (64 FP32 members) only
                     "traffic": (traffic_from_profiles("ensemble_step_kernel_batch_q19_64members")
                                 if a.storage == "f32" and a.members == 64 and not a.small_members else None),
                     "kernel": "step_kernel_batch (graph; faces and inflow planes folded)", "bytes_per_cell": bpc_bytes,
                     "kernel_ms": round(step_ms, 5)},
        "tuned_launch": {"block": block, "blocks_per_cu": bpc, "ms_per_step": round(tune_ms, 5)},
        "per_member_launches": {"value": round(mlups_pm, 1), "unit": "MLUPS", "steps": n_pm,
                                "ms_per_step": round(t_pm / n_pm * 1e3, 5)},
        "speedup_vs_per_member": round(mlups / mlups_pm, 3),
        "cpu_baseline": None, "finite": finite,
    }
    print(json.dumps(line))


def bench_adjoint(a):
    import torch
    from lbm_amd.torch_adjoint import ForcingProblem
    nsteps = a.steps
    print(f"[bench] adjoint problem 192x96x96, {nsteps} steps", file=sys.stderr, flush=True)
    p32 = ForcingProblem(storage="f32", nsteps=nsteps, every=20)
    F = p32.force_field(p32.a0).numpy()
    s = p32.shear_profile(p32.c0).numpy()
    p32.evaluate(F, s, adjoint=False)          # warm-up (allocations, first launches)
    p32.evaluate(F, s, adjoint=True)           # the adjoint's workspace stays with the context
    t0 = time.perf_counter()
    J_fwd = p32.evaluate(F, s, adjoint=False)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    J_adj, gF32, gS32 = p32.evaluate(F, s, adjoint=True)
    t_adj = time.perf_counter() - t0
    print(f"[bench] forward {t_fwd:.3f} s, adjoint run {t_adj:.3f} s", file=sys.stderr, flush=True)
    # parameter gradients through the torch harness
    def grads(prob):
        av = prob.a0.clone().requires_grad_(True)
        cv = prob.c0.clone().requires_grad_(True)
        J = prob.J(av, cv)
        J.backward()
        return float(J.detach()), av.grad.numpy().copy(), cv.grad.numpy().copy()
    J32, ga32, gc32 = grads(p32)
    p32.close()
    p64 = ForcingProblem(storage="f64", nsteps=nsteps, every=20)
    J64, ga64, gc64 = grads(p64)
    which = [("a", 0), ("a", 17), ("a", 40), ("c", 0), ("c", 1), ("c", 5)]
    h = 1e-6
    fd = []
    print("[bench] FP64 central differences", file=sys.stderr, flush=True)
    for kind, i in which:
        ap, am, cp, cm = p64.a0.clone(), p64.a0.clone(), p64.c0.clone(), p64.c0.clone()
        (ap if kind == "a" else cp)[i] += h
        (am if kind == "a" else cm)[i] -= h
        fd.append((p64.J_value(ap, cp) - p64.J_value(am, cm)) / (2 * h))
    p64.close()
    fd = np.array(fd)
    pick = lambda ga, gc: np.array([ga[i] if k == "a" else gc[i] for k, i in which])
    # relative to the largest difference quotient; with too few steps for any signal to reach the
    # objective box every quotient is 0 and the comparison is undefined (reported as null)
    scale = np.abs(fd).max()
    e32 = float(np.abs(pick(ga32, gc32) - fd).max() / scale) if scale > 0 else None
    e64 = float(np.abs(pick(ga64, gc64) - fd).max() / scale) if scale > 0 else None
    # FP64 tangent run along a seeded direction vs the FP64 adjoint's prediction
    print("[bench] FP64 tangent run", file=sys.stderr, flush=True)
    rng = np.random.default_rng(3)
    da = torch.from_numpy(rng.uniform(-1e-5, 1e-5, len(ga64)))
    dc = torch.from_numpy(rng.uniform(-0.1, 0.1, len(gc64)))
    pt = ForcingProblem(storage="f64t", nsteps=nsteps, every=20)
    Jt, dJt = pt.directional(pt.a0, pt.c0, da, dc)
    pt.close()
    pred64 = float(ga64 @ da.numpy() + gc64 @ dc.numpy())
    pred32 = float(ga32 @ da.numpy() + gc32 @ dc.numpy())
    full64 = np.concatenate([ga64, gc64])
    full32 = np.concatenate([ga32, gc32])
    e_full = float(np.abs(full32 - full64).max() / np.abs(full64).max()) if np.abs(full64).max() > 0 else None
    cells = 192 * 96 * 96
    line = {
        "metric": "adjoint MLUPS (D3Q19 reg+Vreman, checkpoint + recompute + reverse)",
        "value": round(cells * nsteps / t_adj / 1e6, 1), "unit": "MLUPS", "n_gpus": 1, "steps": nsteps, "warmup": 1,
        "ms_per_step": round(t_adj / nsteps * 1e3, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded populations, blob centres and amplitudes)",
        "config": {"workload": "D3Q19 regularized + Vreman, 192x96x96, inflow/outflow i, periodic j, no-slip k=1 / "
                               "free-slip k=nz, 64 Gaussian force blobs (sigma 4) near i=40, 8-coefficient power-law "
                               "inflow, objective: time-mean KE over i in [140,180], checkpoints every 20",
                   "cells": cells, "parallelism": "single GPU"},
        "forward_mlups": round(cells * nsteps / t_fwd / 1e6, 1), "adjoint_over_forward": round(t_adj / t_fwd, 3),
        "J": {"f32": J_adj, "f64": J64, "forward_f32": J_fwd},
        "grad_check": {"checked": [f"{k}{i}" for k, i in which], "fd64": fd.tolist(),
                       "adj32": pick(ga32, gc32).tolist(), "adj64": pick(ga64, gc64).tolist(),
                       "max_rel_err_f32_vs_fd64": e32, "max_rel_err_f64_vs_fd64": e64,
                       "max_rel_err_f32_vs_f64_all_72": e_full,
                       "tangent64_dJ": dJt, "adj64_dJ": pred64, "adj32_dJ": pred32,
                       "rel_err_adj64_vs_tangent64": abs(pred64 - dJt) / abs(dJt) if dJt else None,
                       "rel_err_adj32_vs_tangent64": abs(pred32 - dJt) / abs(dJt) if dJt else None},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line))


def bench_tile(a):
    """One GPU's share of the headline workload tiled along j (tile 4 of 8: j = 73..96, through the
    cylinder), with its halo exchange every step: pack both boundary planes, hand each buffer over
    (a device-to-device copy on the context's stream where RCCL would send/recv), unpack.  With
    --split each step is two launches (lbm_step_part), the pack after the first, and the transfer
    runs on a stream of its own during the second."""
    import torch
    from lbm_amd.domain import BND_HALO, PART_EDGE, PART_INNER, Domain
    w = workload(a.storage)
    nx, ny, nz = w["dims"]
    ntile = a.tiles
    if ny % ntile:
        raise SystemExit(f"ny = {ny} is not a multiple of --tiles {ntile}")
    nt = ny // ntile
    t = min(3, ntile - 1)
    j0, j1 = t * nt + 1, (t + 1) * nt
    nrturb = max(1001, a.steps + a.warmup + 1)
    print(f"[bench] tile j = {j0}..{j1} of {w['dims']} storage={a.storage}", file=sys.stderr, flush=True)
    f0, turb, mask = make_inputs(w, nrturb)
    idx = [(j - 1) % ny + 1 for j in range(j0 - 1, j1 + 2)]
    f0t = np.ascontiguousarray(f0[:, idx])
    maskt = np.ascontiguousarray(mask[:, idx])
    turbt = [np.ascontiguousarray(x[:, :, j0 - 1:j1]) for x in turb[:3]]
    del f0, mask
    dom = Domain(w["q"], (nx, nt, nz), a.storage, ibgk=w["ibgk"], ihrr=w["ihrr"], ivreman=w["ivreman"],
                 smagorinsky=w["smagorinsky"], kinevisc=w["kinevisc"], ibnd=w["ibnd"], jbnd=BND_HALO,
                 kbnd=w["kbnd"], iturb_pos=w["ip"])
    dom.set_state(f0t)
    dom.set_inflow(w["uini"], w["udir"])
    dom.set_blanking(maskt)
    dom.set_turbulence(turbt[0], turbt[1], turbt[2], nrturb, w["ampl"])
    nb = dom.halo_bytes()
    snd = torch.empty(2 * nb, dtype=torch.uint8, device="cuda")   # side 0's halo, then side 1's
    rcv = torch.empty(2 * nb, dtype=torch.uint8, device="cuda")
    ext = torch.cuda.ExternalStream(dom.stream_ptr())
    side = torch.cuda.Stream()
    packed, moved = torch.cuda.Event(), torch.cuda.Event()

    def step(it):
        dom.step(it, 1)
        dom.halo_pack(2, snd.data_ptr())
        with torch.cuda.stream(ext):   # the transfer (RCCL send/recv); this tile is its own neighbour
            rcv[:nb].copy_(snd[nb:])
            rcv[nb:].copy_(snd[:nb])
        dom.halo_unpack(2, rcv.data_ptr())

    def step_split(it):
        # step it + 1 (as dom.step(it, 1)): edge rows, pack; the transfer on its own stream during
        # the inner rows; unpack after both
        dom.step_part(it + 1, PART_EDGE)
        dom.halo_pack(2, snd.data_ptr())
        packed.record(ext)
        dom.step_part(it + 1, PART_INNER)
        side.wait_event(packed)
        with torch.cuda.stream(side):
            rcv[:nb].copy_(snd[nb:])
            rcv[nb:].copy_(snd[:nb])
            moved.record(side)
        ext.wait_event(moved)
        dom.halo_unpack(2, rcv.data_ptr())

    if a.split:
        step = step_split

    for it in range(1, a.warmup + 1):
        step(it)
    dom.synchronize()
    dom.set_timing(True, every=10)
    t0 = time.perf_counter()
    for it in range(a.warmup + 1, a.warmup + a.steps + 1):
        step(it)
    dom.synchronize()
    secs = time.perf_counter() - t0
    tm = dom.timing_read()
    rho, u, v, ww = dom.macrovars()
    finite = bool(np.isfinite(rho).all() and np.isfinite(u).all())
    dom.close()
    cells = nx * nt * nz
    # split steps are not bracketed by the step timing
    kern_ms = tm["step_ms"] / tm["n_step"] if tm["n_step"] else None
    bpc = bytes_per_cell(w["q"], a.storage)
    ms = secs / a.steps * 1e3
    print(json.dumps({
        "metric": "MLUPS of one GPU's j-tile (1/%d of 384x192x128) with its halo exchange" % ntile,
        "value": round(cells * a.steps / secs / 1e6, 1), "unit": "MLUPS", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": a.storage, "data": "synthetic (the headline workload's inputs, tile j=%d..%d)" % (j0, j1),
        "config": {"workload": "tile %dx%dx%d + 2 halo planes of %d bytes" % (nx, nt, nz, nb), "storage": a.storage,
                   "parallelism": "single GPU (one tile of a j-decomposition)"},
        "split_steps": bool(a.split),
        "step_kernel_ms": round(kern_ms, 5) if kern_ms else None,
        "step_kernel_GBps": round(bpc * cells / (kern_ms * 1e-3) / 1e9, 1) if kern_ms else None,
        "halo_ms_per_step": round(ms - kern_ms, 5) if kern_ms else None, "finite": finite}), flush=True)


def bench_q19(a):
    """The D3Q19 BGK box physics at the storage-study size: D3Q19 BGK (second-order equilibrium), tau = 0.6
    (nu = 1/30), periodic i and j, no-slip k, 768x256x192, u = (0.05, 0, 0) + U(-1e-3, 1e-3) per
    component (seed 1), no solids, no forcing."""
    from lbm_amd import inputs
    from lbm_amd.domain import Domain
    dims = (768, 256, 192)
    nx, ny, nz = dims
    print(f"[bench] building inputs {dims} q=19 storage={a.storage}", file=sys.stderr, flush=True)
    dt = np.float64 if a.storage == "f64" else np.float32
    f0 = inputs.perturbed_uniform_flow(19, dims, seed=1, u0=(0.05, 0.0, 0.0), order=2, dtype=dt)
    dom = Domain(19, dims, a.storage, ibgk=2, ihrr=0, ivreman=0, kinevisc=0.0333333333, ibnd=0, jbnd=0,
                 kbnd=11, block=tuple(a.block))
    dom.set_state(f0)
    del f0
    print(f"[bench] warmup {a.warmup}", file=sys.stderr, flush=True)
    dom.step(1, a.warmup)
    dom.synchronize()
    dom.set_timing(True, every=10)
    print(f"[bench] timed {a.steps} steps", file=sys.stderr, flush=True)
    ev = StreamEvents(dom.stream_ptr())
    t0 = time.perf_counter()
    ev.record(0)
    dom.step(1 + a.warmup, a.steps)
    ev.record(1)
    dom.synchronize()
    secs = time.perf_counter() - t0
    region_ms = ev.elapsed_ms()
    tm = dom.timing_read()
    rho, u, v, ww = dom.macrovars()
    finite = bool(np.isfinite(rho).all() and np.isfinite(u).all())
    dom.close()
    cells = nx * ny * nz
    sampled_ms = tm["step_ms"] / max(tm["n_step"], 1)
    # events around the timed region (one step-kernel launch per step: the face pass is folded)
    kern_ms = region_ms / a.steps if tm["n_bc"] == 0 else sampled_ms
    bpc = bytes_per_cell(19, a.storage)
    achieved = bpc * cells / (kern_ms * 1e-3) / 1e9
    print(json.dumps({
        "metric": "MLUPS + %HBM roofline (D3Q19 BGK)", "value": round(cells * a.steps / secs / 1e6, 1),
        "unit": "MLUPS", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(secs / a.steps * 1e3, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": a.storage, "data": "synthetic (seeded perturbed uniform flow)",
        "config": {"workload": "D3Q19 BGK tau=0.6, 768x256x192, periodic i and j, no-slip k", "cells": cells,
                   "storage": a.storage, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles("step_kernel_q19_bgk_" + a.storage), "kernel": "step_kernel",
                     "bytes_per_cell": bpc, "kernel_ms": round(kern_ms, 5), "kernel_ms_sampled": round(sampled_ms, 5),
                     "bc_ms_per_step": round(tm["bc_ms"] / max(tm["n_bc"], 1), 5)},
        "cpu_baseline": None, "finite": finite}), flush=True)


def bench_large(a):
    """The headline physics (D3Q27 reg + Vreman, inflow/outflow i, periodic j, no-slip k=1 /
    free-slip k=nz, cylinder r=12 at (96,96), inflow forcing at i=10) at 768x256x192."""
    w = workload(a.storage, (768, 256, 192))
    nx, ny, nz = w["dims"]
    nrturb = max(1001, a.steps + a.warmup + 1)
    print(f"[bench] building inputs {w['dims']} q={w['q']} storage={a.storage}", file=sys.stderr, flush=True)
    f0, turb, mask = make_inputs(w, nrturb)
    res = run_timed(w, a.storage, f0, turb, mask, nrturb, a, info=True)
    cells = nx * ny * nz
    print(json.dumps({
        "metric": "MLUPS + %HBM roofline (D3Q27 reg+Vreman, 768x256x192)", "value": round(res["mlups"], 1),
        "unit": "MLUPS", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(res["ms_per_step"], 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": STORAGE_DTYPE[a.storage],
        "data": "synthetic (seeded perturbed uniform flow, seeded smooth inflow forcing)",
        "config": {"workload": "D3Q27 3rd-order regularized + Vreman (c=0.07), 768x256x192, inflow/outflow i, "
                               "periodic j, no-slip k=1 / free-slip k=nz, cylinder r=12 at (96,96), inflow forcing at "
                               "i=10, nu=1e-4", "cells": cells, "storage": a.storage, "parallelism": "single GPU",
                   "buffer_bytes": res["info"]["device_bytes"] // 2, "row_pitch": res["info"]["row_pitch"]},
        "roofline": roof_of(res, None, None), "finite": res["finite"], "cpu_baseline": None}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: the workload's configuration: headline, tile and q19 1000, "
                         "ensemble 500, adjoint 400)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--storage", default="f32", choices=["f32", "f64", "f16s", "bf16s"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=8, help="CPU baseline sample, all threads")
    ap.add_argument("--cpu-steps-1t", type=int, default=2, help="CPU baseline sample on one thread (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=default_cpu_threads(),
                    help="threads of the CPU baseline (default: the CPUs this process may run on, at most "
                         "OMP_NUM_THREADS when that is set)")
    ap.add_argument("--ramp-ms", type=float, default=300.0,
                    help="untimed steps for this many ms before the warm-up (GPU clock ramp)")
    ap.add_argument("--no-low-precision", action="store_true",
                    help="headline: skip the FP16/BF16-storage timing and the error runs")
    ap.add_argument("--error-steps", type=int, default=1000,
                    help="headline: step count of the storage-error comparison against FP64 (0: skip)")
    ap.add_argument("--block", type=int, nargs=2, default=(0, 0))
    ap.add_argument("--workload", default="headline", choices=["headline", "ensemble", "adjoint", "tile", "q19", "large"])
    ap.add_argument("--tiles", type=int, default=8)
    ap.add_argument("--split", action="store_true",
                    help="tile workload: split steps (edge rows, pack, inner rows during the transfer)")
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--small-members", action="store_true",
                    help="ensemble workload: members of 24-40 x 16-24 x 16-24 cells instead of {96,128,160}x{64,96}x{64,96}")
    a = ap.parse_args()
    if a.steps is None:
        a.steps = {"ensemble": 500, "adjoint": 400}.get(a.workload, 1000)
    if a.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU; --gpus must be 1 (a j-tiled domain "
                                   "measures one GPU's share: --workload tile)"}))
        sys.exit(2)
    if a.workload == "ensemble":
        bench_ensemble(a)
        return
    if a.workload == "adjoint":
        bench_adjoint(a)
        return
    if a.workload == "tile":
        bench_tile(a)
        return
    if a.workload == "q19":
        bench_q19(a)
        return
    if a.workload == "large":
        bench_large(a)
        return

    headline(a)


STORAGE_DTYPE = {"f32": "f32", "f64": "f64", "f16s": "f16-store/f32", "bf16s": "bf16-store/f32"}


class StreamEvents:
    """Two HIP events recorded on the library's own stream through the HIP runtime the library is
    linked against (libamdhip64.so.7, already loaded with it), so that no second runtime copy takes
    part in the measurement."""

    def __init__(self, stream_ptr):
        import ctypes as C
        self.C = C
        self.hip = C.CDLL("libamdhip64.so.7")
        for f, args in (("hipEventCreate", [C.c_void_p]), ("hipEventRecord", [C.c_void_p, C.c_void_p]),
                        ("hipEventSynchronize", [C.c_void_p]), ("hipEventDestroy", [C.c_void_p]),
                        ("hipEventElapsedTime", [C.c_void_p, C.c_void_p, C.c_void_p])):
            getattr(self.hip, f).argtypes = args
            getattr(self.hip, f).restype = C.c_int
        self.stream = C.c_void_p(stream_ptr)
        self.ev = [C.c_void_p(), C.c_void_p()]
        for e in self.ev:
            if self.hip.hipEventCreate(C.byref(e)) != 0:
                raise RuntimeError("hipEventCreate failed")

    def record(self, i):
        if self.hip.hipEventRecord(self.ev[i], self.stream) != 0:
            raise RuntimeError("hipEventRecord failed")

    def elapsed_ms(self):
        ms = self.C.c_float()
        if (self.hip.hipEventSynchronize(self.ev[1]) != 0 or
                self.hip.hipEventElapsedTime(self.C.byref(ms), self.ev[0], self.ev[1]) != 0):
            raise RuntimeError("hipEventElapsedTime failed")
        for e in self.ev:
            self.hip.hipEventDestroy(e)
        return ms.value


def run_timed(w, storage, f0, turb, mask, nrturb, a, info=False):
    """One storage form of the headline workload: untimed clock ramp and warm-up, then exactly
    a.steps timed steps.  Two HIP events on the library's stream bracket the whole timed region
    (one step-kernel launch per step when the boundary passes are folded: the region's time per
    step is then the kernel's average launch duration, inter-launch gaps included); HIP events also
    bracket the step kernel of every max(1, steps // 10)-th timed step (at least 10 samples from 10
    steps on; each record costs the stream a few microseconds); the wall clock brackets the whole
    timed region."""
    from lbm_amd.domain import Domain
    dom = Domain(w["q"], w["dims"], storage, ibgk=w["ibgk"], ihrr=w["ihrr"], ivreman=w["ivreman"],
                 smagorinsky=w["smagorinsky"], kinevisc=w["kinevisc"], ibnd=w["ibnd"], jbnd=w["jbnd"],
                 kbnd=w["kbnd"], iturb_pos=w["ip"], block=tuple(a.block))
    dom.set_state(f0)
    dom.set_inflow(w["uini"], w["udir"])
    dom.set_blanking(mask)
    dom.set_turbulence(turb[0], turb[1], turb[2], nrturb, w["ampl"])
    it = 1
    # clock ramp (untimed): step until a.ramp_ms of wall time has passed
    t0 = time.perf_counter()
    while (time.perf_counter() - t0) * 1e3 < a.ramp_ms:
        dom.step(it, 10)
        it += 10
        dom.synchronize()
    print(f"[bench] {storage}: ramp {it - 1} steps, warmup {a.warmup}", file=sys.stderr, flush=True)
    dom.step(it, a.warmup)
    it += a.warmup
    dom.synchronize()
    every = max(1, a.steps // 10)
    dom.set_timing(True, every=every)
    print(f"[bench] {storage}: timed {a.steps} steps (kernel events every {every})", file=sys.stderr, flush=True)
    ev = StreamEvents(dom.stream_ptr())
    t0 = time.perf_counter()
    ev.record(0)
    dom.step(it, a.steps)
    ev.record(1)
    dom.synchronize()
    secs = time.perf_counter() - t0
    region_ms = ev.elapsed_ms()
    tm = dom.timing_read()
    rho, u, v, ww = dom.macrovars()
    finite = bool(np.isfinite(rho).all() and np.isfinite(u).all())
    layout = dom.info() if info else None
    dom.close()
    cells = w["dims"][0] * w["dims"][1] * w["dims"][2]
    bpc = bytes_per_cell(w["q"], storage)
    sampled_ms = tm["step_ms"] / max(tm["n_step"], 1)
    region_step_ms = region_ms / a.steps
    # the region's time per step is the step kernel's when no boundary kernel runs beside it
    one_launch = tm["n_bc"] == 0
    kern_ms = region_step_ms if one_launch else sampled_ms
    ms = secs / a.steps * 1e3
    return {"info": layout, "mlups": cells * a.steps / secs / 1e6, "ms_per_step": ms, "kernel_ms": kern_ms,
            "kernel_timing": ("HIP events around the timed region on the library's stream (one step-kernel "
                              "launch per step)" if one_launch else "HIP events around sampled step kernels"),
            "kernel_ms_sampled": sampled_ms, "region_ms_per_step": region_step_ms,
            "kernel_samples": tm["n_step"], "bc_ms": tm["bc_ms"] / max(tm["n_bc"], 1),
            "achieved": bpc * cells / (kern_ms * 1e-3) / 1e9,
            "achieved_wall": bpc * cells / (ms * 1e-3) / 1e9, "bpc": bpc, "finite": finite}


def storage_errors(w, storages, checkpoints, seed):
    """Rel-L2 errors of each storage form against FP64 storage (FP64 arithmetic) on the headline
    workload, every run from the equilibrium of the same FP64 moments."""
    from lbm_amd import inputs, precision_study
    dims = w["dims"]
    q, opts, cyl = precision_study.physics("regvre", dims)
    assert q == w["q"] and cyl == w["cyl"] and opts["kinevisc"] == w["kinevisc"]
    nrturb = 250
    moments = inputs.perturbed_moments(dims, seed, dtype=np.float64)
    turb = inputs.turbulence_fields(seed, dims[1], dims[2], nrturb, corr=w["corr"], dtype=np.float64)
    ref = precision_study.run("f64", q, opts, dims, cyl, moments, turb, checkpoints, nrturb)
    out = {}
    for st in storages:
        x = precision_study.run(st, q, opts, dims, cyl, moments, turb, checkpoints, nrturb)
        out[st] = {str(n): precision_study.compare(x[n], ref[n]) for n in checkpoints}
        out[st]["finite"] = bool(all(np.isfinite(x[n][1]).all() for n in checkpoints))
        del x
    return out


def roof_of(r, traffic, traffic_src):
    return {"bound": "hbm", "achieved": round(r["achieved"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(r["achieved"] / HBM_PEAK_GBS, 4),
            "frac_wall": round(r["achieved_wall"] / HBM_PEAK_GBS, 4),
            "traffic": traffic, "traffic_source": traffic_src,
            "kernel": "step_kernel", "bytes_per_cell": r["bpc"], "kernel_ms": round(r["kernel_ms"], 5),
            "kernel_timing": r["kernel_timing"], "kernel_ms_sampled": round(r["kernel_ms_sampled"], 5),
            "kernel_samples": r["kernel_samples"], "bc_ms_per_step": round(r["bc_ms"], 5)}


def headline(a):
    w = workload(a.storage)
    nx, ny, nz = w["dims"]
    # turbulence slices for every step index the run reaches (ramp included)
    # one turbulence slice per step of the warm-up and timed region (the ramp's steps wrap,
    # lit = mod(it, nrturb) as in src/m_inflow_turbulence_forcing.F90:45-46)
    nrturb = max(1001, a.steps + a.warmup + 1)
    print(f"[bench] building inputs {w['dims']} q={w['q']} storage={a.storage}", file=sys.stderr, flush=True)
    f0, turb, mask = make_inputs(w, nrturb)
    cells = nx * ny * nz
    res = run_timed(w, a.storage, f0, turb, mask, nrturb, a)
    kname = "step_kernel_q27_reg_vreman_" + a.storage
    line = {
        "metric": "MLUPS + %HBM roofline (D3Q27 reg+Vreman), FP16-store vs FP32 at stated error",
        "value": round(res["mlups"], 1), "unit": "MLUPS", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(res["ms_per_step"], 5), "higher_is_better": True, "scaling": "weak",
        # the reference publishes no number for this metric (its README times a laptop GPU on
        # another grid with unstated collision flags)
        "vs_baseline": None, "dtype": STORAGE_DTYPE[a.storage],
        "data": "synthetic (seeded perturbed uniform flow, seeded smooth inflow forcing)",
        "config": {"workload": "D3Q27 3rd-order regularized + Vreman (c=0.07), 384x192x128, inflow/outflow i, "
                               "periodic j, no-slip k=1 / free-slip k=nz, cylinder r=12 at (96,96), "
                               "inflow forcing at i=10, nu=1e-4",
                   "cells": cells, "storage": a.storage, "parallelism": "single GPU"},
        "roofline": roof_of(res, traffic_from_profiles(kname),
                            f"profiles/{kname}.json (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; copy-kernel scaled)"),
        "finite": res["finite"],
    }
    if a.storage == "f32" and not a.no_low_precision:
        # the FP16 half of the metric: the same workload with weight-shifted FP16 (and BF16) storage
        lp = {}
        for st in ("f16s", "bf16s"):
            r = run_timed(w, st, f0, turb, mask, nrturb, a)
            kn = "step_kernel_q27_reg_vreman_" + st
            lp[st] = {"value": round(r["mlups"], 1), "unit": "MLUPS", "ms_per_step": round(r["ms_per_step"], 5),
                      "ratio_vs_f32": round(r["mlups"] / res["mlups"], 3),
                      "roofline": roof_of(r, traffic_from_profiles(kn),
                                          f"profiles/{kn}.json (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)"),
                      "finite": r["finite"]}
        if a.error_steps:
            cps = sorted(set([100] + [a.error_steps]))
            print(f"[bench] storage errors vs FP64 after {cps} steps", file=sys.stderr, flush=True)
            errs = storage_errors(w, ("f32", "f16s", "bf16s"), cps, seed=w["seed"])
            for st in ("f16s", "bf16s"):
                lp[st]["error_vs_f64"] = errs[st]
            line["error_vs_f64"] = errs["f32"]
            line["error_method"] = (
                "relative L2 over interior cells of the FP64-storage (FP64 arithmetic) run: 'velocity' is "
                "||(u,v,w) - (u,v,w)_64|| / ||(u,v,w)_64||; per field rho-1, u, v, w; mean_profile and "
                "tke_profile are the max deviations of the i,j-mean u(k) and TKE(k) relative to their maxima; "
                "every run starts from the equilibrium of the same seeded FP64 moments")
        line["f16_storage"] = lp["f16s"]
        line["bf16_storage"] = lp["bf16s"]
    cpu = None
    if not a.no_cpu_baseline and a.storage == "f32":
        print(f"[bench] cpu baseline: reference OpenMP build, {a.cpu_steps} steps", file=sys.stderr, flush=True)
        try:
            cpu = cpu_baseline(w, f0, turb, a.cpu_steps, a.cpu_threads, a.cpu_steps_1t)
        except Exception as e:  # the baseline is reported, never required
            print(f"[bench] cpu baseline failed: {e}", file=sys.stderr)
    line["cpu_baseline"] = cpu
    print(json.dumps(line))


if __name__ == "__main__":
    main()
