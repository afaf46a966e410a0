"""Host-side parts of the turbine forcing, no device: the library's conversion of the infile values
(lbm_amd.turbines.from_infile) equals the oracle's restatement of turbine_initialize, which the
reference-build goldens pin bit for bit, and the airfoil reader returns the tables those goldens
were made with."""
import os

import numpy as np
import pytest

import cases
import oracle
from lbm_amd import turbines

AIRFOILS = "/root/reference/airfoils"


@pytest.mark.parametrize("name", sorted(cases.TURBINES))
def test_infile_conversion_matches_the_oracle(name):
    meta, fx = cases.load(name)
    t = cases.TURBINES[name]
    foils = [fx["foil"][k, :n] for k, n in enumerate(fx["nang"])]
    mine = turbines.from_infile(t["pos"], t["rpm"], t["pitch"], cases.TURBINE_LENGTH, cases.TURBINE_VEL, foils,
                                np.float32, theta=t["theta"])
    ref = oracle.turbines_from_infile(t["pos"], t["rpm"], t["pitch"], cases.TURBINE_LENGTH, cases.TURBINE_VEL,
                                      foils, "f32", theta=t["theta"])
    assert np.array_equal(mine["hub"][:, 0], ref["xhub"]) and np.array_equal(mine["hub"][:, 2], ref["zhub"])
    for k in ("yaw", "tilt", "omegand", "pitch", "theta", "relm", "dc", "chord", "twist", "nfoil"):
        assert np.array_equal(mine[k], ref[k]), k
    # yaw/tilt in radians, omegand = 2 pi rpm / 60 * C_l / C_u
    assert np.isclose(mine["omegand"][0], 2 * np.pi * t["rpm"] / 60 * cases.TURBINE_LENGTH / cases.TURBINE_VEL,
                      rtol=1e-6)
    assert np.allclose(mine["yaw"], np.radians([p[3] for p in t["pos"]]), rtol=1e-6)


@pytest.mark.skipif(not os.path.isdir(AIRFOILS), reason="reference airfoil tables not present")
def test_airfoil_reader_matches_the_fixture_tables():
    meta, fx = cases.load("q27_regvre_turbine")
    tabs = turbines.read_airfoils(AIRFOILS)
    assert [len(a) for a in tabs] == list(fx["nang"])
    for k, a in enumerate(tabs):
        assert np.array_equal(a, fx["foil"][k, :len(a)])
    # a table is ordered in angle and spans the full circle
    assert all(np.all(np.diff(a[:, 0]) > 0) and a[0, 0] == -180 and a[-1, 0] == 180 for a in tabs if len(a) > 3)
