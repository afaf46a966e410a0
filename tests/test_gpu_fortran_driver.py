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


def write_airfoils(directory, fx):
    """The fixture's airfoil tables as ./Airfoils/*.dat in the layout src/m_nrelreadfoil.F90:33-47
    reads: 50 header lines, the row count, 2 lines, rows "alpha cl cd cm" (values printed exactly)."""
    from lbm_amd.turbines import FOIL_FILES
    os.makedirs(directory, exist_ok=True)
    for k, fn in enumerate(FOIL_FILES):
        n = int(fx["nang"][k])
        rows = ["! header line %d" % (r + 1) for r in range(50)]
        rows += [f"        {n}   NumAlf", "!    Alpha      Cl      Cd        Cm", "!    (deg)      (-)     (-)       (-)"]
        rows += ["  %.9g  %.9g  %.9g  0.0" % tuple(float(x) for x in fx["foil"][k, r]) for r in range(n)]
        with open(os.path.join(directory, fn), "w") as fh:
            fh.write("\n".join(rows) + "\n")


@pytest.mark.parametrize("name", ["q27_regvre_inflow_cyl", "q27_regvre_inflow_dp", "q19_bgk_inflow_dir",
                                  "q27_regvre_closed", "q27_bgk3_inflow_series", "q27_regvre_turbine",
                                  "q19_bgk_turbines2"])
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
    for fname, text in cases.case_files(name).items():   # uvel_shear.dat, uvel_time.dat
        (tmp_path / fname).write_text(text)
    if name in cases.TURBINES:
        opts.update(cases.infile_units(name))
        write_airfoils(str(tmp_path / "Airfoils"), fx)
        fio.write_theta(str(tmp_path / "restart" / fio.restart_name(nt0, prefix="theta")),
                        np.array(cases.TURBINES[name]["theta"], np.float32))
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
    if name in cases.TURBINES:
        th = fio.read_theta(str(tmp_path / "restart" / fio.restart_name(nt1, prefix="theta")))
        assert np.array_equal(th, fx["theta_out"]), (th, fx["theta_out"])
    # the testing file (src/m_testing.F90) is written on the first run and compared on a second
    assert (tmp_path / "testing" / fio.restart_name(nt1, prefix="testing")).exists()


def test_fortran_driver_cold_start_generated_turbulence_and_averages(tmp_path):
    """nt0 = 0: equilibrium start from the inflow profile, turbulence fields generated on the
    device at start-up and every nrturb steps, time averages over avestart < it <= avesave."""
    exe = os.path.join(BUILD, "lbm_driver_f32")
    if not os.path.exists(exe):
        pytest.skip("Fortran driver not built (no amdflang)")
    nx, ny, nz, nrturb = 40, 16, 12, 8
    txt = oracle.infile_text(nt0=0, nt1=20, turb=True, ampl=1e-3, nrturb=nrturb, ibgk=3, ihrr=1, ivreman=1,
                             smagorinsky=0.16733, ibnd=1, jbnd=0, kbnd=12, uini=0.05, visc=1e-3)
    assert " F F\n 20000\n 40000\n" in txt
    (tmp_path / "infile.in").write_text(txt.replace(" F F\n 20000\n 40000\n", " T F\n 5\n 15\n"))
    print(f"driver cold start {nx}x{ny}x{nz}, nrturb={nrturb}", flush=True)
    p = subprocess.run([exe, str(nx), str(ny), str(nz), "27"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    print(p.stdout[-2000:], p.stderr[-1000:], flush=True)
    assert p.returncode == 0
    f = fio.read_restart(str(tmp_path / "restart" / fio.restart_name(20)))
    assert f.shape == (nz + 2, ny + 2, nx + 2, 27) and np.isfinite(f).all()
    uu, vv, ww, rr = fio.read_turbulence(str(tmp_path / "restart" / fio.restart_name(20, prefix="turbulence")))
    assert uu.shape == (nrturb + 1, nz, ny)
    assert np.abs(uu[1:].mean(axis=(1, 2))).max() < 1e-5
    assert np.abs(uu[1:].std(axis=(1, 2), ddof=1) - 1).max() < 1e-4
    assert "Averages of     10 steps" in p.stdout
    assert (tmp_path / "averages" / "averages_0000_000015.uf").exists()
