"""Actuator-line turbines for Domain.set_turbines (lbm_set_turbines in include/lbm_hip.h).

The reference configures its rotors in infile.in (src/m_readinfile.F90:143-161), takes the blade
of the NREL 5-MW turbine (src/mod_nrel5mw.F90) and reads eight airfoil tables from ./Airfoils
(src/m_nrelreadfoil.F90:8-57); turbine_initialize (src/m_turbine_initialize.F90:29-70) converts to
lattice units.  ``from_infile`` does the same conversions, rounding each step in the REAL kind the
run computes in, and returns the description the library takes.
"""
from __future__ import annotations

import os

import numpy as np

# NREL 5-MW blade sections (published by NREL; src/mod_nrel5mw.F90:24-48): radial position, segment
# length and chord [m], twist [deg], airfoil table (1-based into FOIL_FILES); rotor and hub radius [m].
ROTOR_RADIUS, HUB_RADIUS = 62.940, 1.499
RELM = (2.86670, 5.60000, 8.33330, 11.75000, 15.85000, 19.95000, 24.05000, 28.15000, 32.25000, 36.35000,
        40.45000, 44.55000, 48.65000, 52.75000, 56.16670, 58.90000, 61.63330)
DC = (2.73330,) * 3 + (4.10000,) * 11 + (2.73330,) * 3
CHORD = (3.54200, 3.85400, 4.16700, 4.55700, 4.65200, 4.45800, 4.24900, 4.00700, 3.74800, 3.50200, 3.25600,
         3.01000, 2.76400, 2.51800, 2.31300, 2.08600, 1.41900)
TWIST = (13.3080, 13.3080, 13.3080, 13.3080, 11.4800, 10.1620, 9.0110, 7.7950, 6.5440, 5.3610, 4.1880, 3.1250,
         2.3190, 1.5260, 0.8630, 0.3700, 0.1060)
NFOIL = (1, 1, 2, 3, 4, 4, 5, 6, 6, 7, 7, 8, 8, 8, 8, 8, 8)
FOIL_FILES = ("Cylinder1.dat", "Cylinder2.dat", "DU40_A17.dat", "DU35_A17.dat", "DU30_A17.dat", "DU25_A17.dat",
              "DU21_A17.dat", "NACA64_A17.dat")


def read_airfoil(path: str, dtype=np.float32) -> np.ndarray:
    """One airfoil table in the AeroDyn format the reference reads: 50 header lines, the number of
    rows, 2 more lines, then rows "alpha cl cd cm".  Returns an (n, 3) array: alpha [deg], cl, cd."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    n = int(lines[50].split()[0])
    R = np.dtype(dtype).type
    return np.array([[R(x) for x in lines[53 + r].replace(",", " ").split()[:3]] for r in range(n)], dtype)


def read_airfoils(directory: str, dtype=np.float32):
    """The eight tables of FOIL_FILES from ``directory`` (the reference's ./Airfoils)."""
    return [read_airfoil(os.path.join(directory, f), dtype) for f in FOIL_FILES]


def from_infile(pos, rpm, pitch, length, vel, foils, dtype=np.float32, theta=None, nblades=3):
    """The description Domain.set_turbines takes, from the reference's infile values:

    pos     [(ipos, jpos, kpos, yaw [deg], tilt [deg]), ...] per rotor
    rpm     rotor speed [1/min];  pitch  collective pitch [deg]
    length  C_l, metres per cell;  vel  C_u, m/s per lattice speed (p2l%time = length / vel)
    foils   the airfoil tables (read_airfoils)
    theta   the rotors' azimuths [rad] of the state (restart record), zeros by default
    """
    R = np.dtype(dtype).type
    pi = R(3.14159265358979323846)
    pi2 = R(2.0) * pi
    length, vel = R(length), R(vel)
    omegand = (pi2 * R(rpm) / R(60.0)) * (length / vel)
    n = len(pos)
    return dict(
        hub=np.array([[R(p[0]), R(p[1]), R(p[2])] for p in pos], dtype),
        yaw=np.array([(R(p[3]) / R(360.0)) * pi2 for p in pos], dtype),
        tilt=np.array([(R(p[4]) / R(360.0)) * pi2 for p in pos], dtype),
        omegand=np.full(n, omegand, dtype), pitch=np.full(n, R(pitch), dtype),
        theta=np.zeros(n, dtype) if theta is None else np.asarray(theta, dtype),
        relm=np.array([R(x) / length for x in RELM], dtype), dc=np.array([R(x) / length for x in DC], dtype),
        chord=np.array([R(x) / length for x in CHORD], dtype), twist=np.array(TWIST, dtype),
        nfoil=np.array(NFOIL, np.int32), foils=[np.asarray(f, dtype) for f in foils], nblades=int(nblades))
