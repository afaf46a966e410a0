#!/usr/bin/env python3
"""Writes tests/golden/long_run_stats.npz: the long-run statistics of long_run.py from the
REFERENCE's own builds (FP32 and FP64, amdflang with OpenMP, oracle/build_ref.py), restarted every
SEGMENT steps.  Needs /root/reference; run from the repository root:
    python tests/golden/make_long_run.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

import build_ref  # noqa: E402
import long_run as LR  # noqa: E402
import oracle  # noqa: E402


def run(prec):
    exe = build_ref.build(LR.Q, prec, LR.DIMS, omp=True)
    f = LR.initial_state(np.float32 if prec == "sp" else np.float64)
    solid = LR.blanking()[1:-1, 1:-1, 1:-1].astype(bool)
    samples = []
    wanted = set(LR.sample_steps())
    for n0 in range(0, LR.STEPS, LR.SEGMENT):
        f, _ = oracle.run_reference(exe, f, nt0=1 + n0, nt1=1 + n0 + LR.SEGMENT, experiment="cylinder", threads=8,
                                    **LR.OPTS)
        if n0 + LR.SEGMENT in wanted:
            samples.append(LR.velocities(f))
        print(prec, n0 + LR.SEGMENT, flush=True)
    return LR.statistics(samples, solid)


if __name__ == "__main__":
    out = {}
    for prec in ("sp", "dp"):
        st = run(prec)
        for k, v in st.items():
            out[f"{prec}_{k}"] = v
    meta = dict(q=LR.Q, dims=list(LR.DIMS), steps=LR.STEPS, segment=LR.SEGMENT, seed=LR.SEED, options=LR.OPTS,
                samples=LR.sample_steps(), producer="reference builds (amdflang -fopenmp) via oracle/build_ref.py")
    np.savez_compressed(os.path.join(HERE, "long_run_stats.npz"), meta=np.array(json.dumps(meta)), **out)
    print("written", os.path.getsize(os.path.join(HERE, "long_run_stats.npz")), "bytes")
