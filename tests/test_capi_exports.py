"""The C-ABI library loads and exports exactly the functions include/lbm_hip.h declares.
No device calls (runs without a GPU)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lbm_hip.h")
LIB = os.path.join(ROOT, "lbm_amd", "liblbm_hip.so")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_[a-z_0-9]+)\s*\(", txt)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("liblbm_hip.so not built (python -m lbm_amd.build)")
    import ctypes
    return ctypes.CDLL(LIB)


def test_every_declared_symbol_is_exported(lib):
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), f"{n} declared in lbm_hip.h but not exported"


def test_no_undeclared_lbm_exports():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = sorted({l.split()[-1] for l in out.splitlines() if l.split()[-1].startswith("lbm_")})
    assert exported == declared()


def test_python_binding_matches_header():
    from lbm_amd import domain
    assert sorted(domain.EXPORTS) == declared()


def test_create_without_device_fails_loudly():
    """On a machine without a HIP device lbm_create must return an error, never fall back."""
    import ctypes
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    from lbm_amd.domain import Domain, LBMError
    with pytest.raises(LBMError, match="lbm_create failed"):
        Domain(19, (8, 8, 8))


def test_invalid_descriptor_is_rejected_before_device_use():
    from lbm_amd.domain import Domain, LBMError
    with pytest.raises(LBMError, match="q must be 19 or 27"):
        Domain(15, (8, 8, 8))
    with pytest.raises(LBMError, match="nx, ny, nz"):
        Domain(19, (1, 8, 8))
    with pytest.raises(LBMError, match="boundary codes"):
        Domain(19, (8, 8, 8), kbnd=1)
