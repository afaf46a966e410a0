"""Buffer layouts (lbm_amd/csrc/layout.h), checked on the CPU by lbm_amd/_build/layout_check: every
(population, cell) of the padded grid gets its own element inside the buffer in both the
row-blocked and the plane arrangement, rows start 128-byte aligned, per-cell indices are distinct,
row-blocked buffers are used while element offsets fit 31 bits and a slab of the slowest axis fits
4 GiB (the 5.2 GB buffer of D3Q27 FP32 at 768x256x192 included) with every element of a slab inside its slab's
range; the two launches of a split step cover each cell
of the halo axis once (one or two edge rows per side, tile heights up to 1024, block heights 1, 2,
4, 8)."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "lbm_amd", "_build", "layout_check")


def test_layouts_are_bijective_aligned_and_bounded():
    if not os.path.exists(EXE):
        pytest.fail("layout_check not built (python -m lbm_amd.build)")
    p = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    rows = [json.loads(l) for l in p.stdout.splitlines() if l.strip()]
    assert p.returncode == 0, rows[-1] if rows else p.stderr
    summary = rows[-1]
    assert summary["cases"] == 6 * 4 * 2 * 5 * 2 * 4
    assert summary["rowblock_768x256x192_f32"] == 1 and summary["rowblock_384x192x128_f32"] == 1
    assert summary["rowblock_2048x1024x1024_f16"] == 0
    assert summary["split_heights"] == ((1024 - 3 + 1) + (1024 - 5 + 1)) * 4
    cases = rows[:-1]
    assert any(r["rowblock"] for r in cases) and any(not r["rowblock"] for r in cases)
