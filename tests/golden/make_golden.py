#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/cases.py by running the REFERENCE's own build
(amdflang, oracle/build_ref.py) on the seeded inputs.  Needs /root/reference; run from the
repository root:  python tests/golden/make_golden.py [case ...]
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
import oracle  # noqa: E402
from cases import (BUOYANCY, CASES, TURBINES, case_files, case_inputs, fixture_path, infile_units,  # noqa: E402
                   initial_pottemp, sha)

AIRFOILS = "/root/reference/airfoils"   # the reference's airfoil tables, read at generation time only


def make(name):
    q, prec, dims, steps, seed, exp, nrturb, o = CASES[name]
    f0, turb = case_inputs(name)
    exe = build_ref.build(q, prec, dims)
    opts = dict(o)
    if nrturb:
        opts["nrturb"] = nrturb
        opts.setdefault("ampl", 0.0)
    files = dict(case_files(name))
    theta, foils = None, None
    if name in TURBINES:
        opts.update(infile_units(name))
        theta = np.array(TURBINES[name]["theta"], np.float32 if prec == "sp" else np.float64)
        texts = [open(os.path.join(AIRFOILS, fn)).read() for fn in oracle.NREL5MW_FOILS]
        files.update({"Airfoils/" + fn: t for fn, t in zip(oracle.NREL5MW_FOILS, texts)})
        foils = [oracle.read_foil_text(t, "f32" if prec == "sp" else "f64") for t in texts]
    extra = {}
    if name in BUOYANCY:
        opts.update(infile_units(name))
        # the potential temperature of every step: the reference run for 1, 2, .., steps steps from the
        # same restart writes the field its own transport (heatflux + advection) produced
        pot = [initial_pottemp(name)]
        for n in range(1, steps):
            _, pt, _ = oracle.run_reference(exe, f0, nt0=1, nt1=1 + n, experiment=exp, files=files,
                                            pottemp=pot[0], **opts)
            pot.append(pt)
        res = oracle.run_reference(exe, f0, nt0=1, nt1=1 + steps, experiment=exp, files=files, pottemp=pot[0],
                                   **opts)
        extra = dict(pottemp=np.stack(pot), pottemp_out=res[1])
    else:
        res = oracle.run_reference(exe, f0, nt0=1, nt1=1 + steps, turb_fields=turb, experiment=exp, files=files,
                                   theta=theta, **opts)
    f_out = res[0]
    meta = dict(name=name, q=q, prec=prec, dims=list(dims), steps=steps, seed=seed, experiment=exp,
                nrturb=nrturb, options=o, nt0=1, input_sha256=sha(f0), files=case_files(name),
                turbines=TURBINES.get(name), infile_units=infile_units(name), buoyancy=BUOYANCY.get(name),
                producer="reference build (amdflang) via oracle/build_ref.py")
    arrays = dict(meta=np.array(json.dumps(meta)), f_out=f_out, **extra)
    if turb is not None:
        arrays.update(uu=turb[0], vv=turb[1], ww=turb[2])
    if foils is not None:
        # the airfoil tables as read (deg, cl, cd per row), padded; the azimuths after the run
        nmax = max(len(f) for f in foils)
        tab = np.zeros((len(foils), nmax, 3), foils[0].dtype)
        for k, f in enumerate(foils):
            tab[k, :len(f)] = f
        arrays.update(foil=tab, nang=np.array([len(f) for f in foils], np.int32), theta_out=res[1])
    np.savez_compressed(fixture_path(name), **arrays)
    print(f"{name}: {os.path.getsize(fixture_path(name)) / 1024:.0f} KiB")


if __name__ == "__main__":
    for n in (sys.argv[1:] or CASES):
        make(n)
