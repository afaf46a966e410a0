"""The Fortran driver (lbm_amd/fortran/lbm_driver.F90, bind(C) over the C-ABI) run as the
reference's executable is run: infile.in and restart/restart_0000_<nt0>.uf (plus the turbulence
restart file) in a working directory, restart/restart_0000_<nt1>.uf out.  The output record
must match the reference build's golden vector for the same case to the tolerance of
test_gpu_golden.py (FP32 2e-6 absolute, FP64 1e-12).
"""
import os
import subprocess

import numpy as np
import pytest

import cases
import oracle
from lbm_amd import fortran_io as fio

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "lbm_amd", "_build")


@pytest.mark.parametrize("name", ["q27_regvre_inflow_cyl", "q27_regvre_inflow_dp", "q19_bgk_inflow_dir",
                                  "q27_regvre_closed"])
def test_fortran_driver_matches_reference_build(name, tmp_path):
    q, prec, dims, steps, seed, exp, nrturb, o = cases.CASES[name]
    exe = os.path.join(BUILD, "lbm_driver_" + ("f32" if prec == "sp" else "f64"))
    if not os.path.exists(exe):
        pytest.skip("Fortran driver not built (no amdflang)")
    meta, fx = cases.load(name)
    f0, turb = cases.case_inputs(name)
    nt0, nt1 = 1, 1 + steps
    os.makedirs(tmp_path / "restart")
    fio.write_restart(str(tmp_path / "restart" / fio.restart_name(nt0)), f0)
    opts = dict(o)
    if nrturb:
        opts["nrturb"] = nrturb
        opts.setdefault("ampl", 0.0)
        fio.write_turbulence(str(tmp_path / "restart" / fio.restart_name(nt0, prefix="turbulence")), *turb)
    (tmp_path / "infile.in").write_text(
        oracle.infile_text(nt0=nt0, nt1=nt1, turb=bool(nrturb), experiment=exp, **opts))
    print(f"driver {os.path.basename(exe)} case {name} dims={dims} steps={steps}", flush=True)
    p = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(p.stdout[-1500:], p.stderr[-1500:], flush=True)
    assert p.returncode == 0
    f = fio.read_restart(str(tmp_path / "restart" / fio.restart_name(nt1)))
    ref = fx["f_out"]
    assert f.shape == ref.shape and f.dtype == ref.dtype
    d = np.abs(f.astype(np.float64) - ref.astype(np.float64))
    tol = 2e-6 if ref.dtype == np.float32 else 1e-12
    print(f"max |df| vs reference build {d.max():.3e}", flush=True)
    assert np.isfinite(f).all() and d.max() <= tol
    # the testing file (src/m_testing.F90) is written on the first run and compared on a second
    assert (tmp_path / "testing" / fio.restart_name(nt1, prefix="testing")).exists()
