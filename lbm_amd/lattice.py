"""D3Q19 and D3Q27 lattice tables.

The direction order, weights and opposite pairs are those of src/mod_D3Q27setup.F90:28-35
(D3Q19) and :40-48 (D3Q27).  The order is part of the restart-file layout (f(nl,...)), so it
is kept exactly; opposite directions are adjacent pairs (2,3),(4,5),... in 1-based numbering.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

_CX27 = [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, 0, 0, -1, 1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1]
_CY27 = [0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 0, 0, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1]
_CZ27 = [0, 0, 0, 0, 0, -1, 1, 0, 0, 0, 0, -1, 1, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1, -1, 1]


@dataclass(frozen=True)
class Lattice:
    q: int
    cx: np.ndarray
    cy: np.ndarray
    cz: np.ndarray
    w: np.ndarray
    opposite: np.ndarray

    @property
    def c(self) -> np.ndarray:
        return np.stack([self.cx, self.cy, self.cz], axis=1)


def lattice(q: int) -> Lattice:
    if q not in (19, 27):
        raise ValueError("q must be 19 or 27")
    cx = np.array(_CX27[:q], dtype=np.int64)
    cy = np.array(_CY27[:q], dtype=np.int64)
    cz = np.array(_CZ27[:q], dtype=np.int64)
    n2 = cx * cx + cy * cy + cz * cz
    if q == 19:
        w = np.where(n2 == 0, 1.0 / 3.0, np.where(n2 == 1, 1.0 / 18.0, 1.0 / 36.0))
    else:
        w = np.select([n2 == 0, n2 == 1, n2 == 2], [8.0 / 27.0, 2.0 / 27.0, 1.0 / 54.0], 1.0 / 216.0)
    opp = np.array([0] + [l + 1 if l % 2 == 1 else l - 1 for l in range(1, q)], dtype=np.int64)
    return Lattice(q, cx, cy, cz, w, opp)
