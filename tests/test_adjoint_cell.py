"""The per-cell collision adjoint (collide_adjoint, lbm_amd/csrc/adjoint.h) against central
differences of the step kernel's own collision (collide_cell), both run on the CPU by
lbm_amd/_build/adjoint_check: D3Q19 and D3Q27, BGK / regularized, with and without Vreman, the
third-order equilibrium, body force, solid swap and the kinetic-energy objective term.
Directional derivatives must agree to 1e-7 (double) and 1e-4 (float) relative."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "lbm_amd", "_build", "adjoint_check")


def test_collision_adjoint_matches_central_differences():
    if not os.path.exists(EXE):
        pytest.fail("adjoint_check not built (python -m lbm_amd.build)")
    out = subprocess.run([EXE], capture_output=True, text=True, check=True, timeout=120).stdout
    rows = [json.loads(l) for l in out.splitlines() if l.strip()]
    assert len(rows) == 96
    combos = {(r["q"], r["bgk3"], r["hrr"], r["vre"], r["force"], r["solid"]) for r in rows}
    assert len(combos) == 32
    worst64 = max(r["rel64"] for r in rows)
    worst32 = max(r["rel32"] for r in rows)
    print(f"max relative error: double {worst64:.2e}, float {worst32:.2e}")
    assert worst64 <= 1e-7
    assert worst32 <= 1e-4
