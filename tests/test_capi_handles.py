"""Every export rejects a handle of the wrong kind (an ensemble handle where a context is
expected, and the reverse), a destroyed handle and a null handle with LBM_EINVAL, before it
reads anything beyond the handle's tag: no device call happens, so this runs without a GPU.
The handles here are host buffers that carry the other kind's tag, or none."""
import ctypes as C
import os
import re
import struct

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lbm_hip.h")
LBM_EINVAL = 1
CTX_TAG, ENS_TAG = 0x4c424d43, 0x4c424d45


def exports_taking(handle_type):
    """(name, number of parameters) of the functions whose first parameter is handle_type*."""
    txt = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = []
    for m in re.finditer(r"\b(?:int|void\*|const char\*)\s+(lbm_[a-z_0-9]+)\s*\(([^)]*)\)", txt):
        params = [p.strip() for p in m.group(2).split(",")]
        if re.match(rf"^(const\s+)?{handle_type}\s*\*\s*\w+$", params[0]):
            out.append((m.group(1), len(params)))
    return out


@pytest.fixture(scope="module")
def lib():
    from lbm_amd.domain import load_library
    return load_library()


def fake_handle(tag):
    """A zeroed host block of 64 KiB whose first 4 bytes are tag."""
    buf = C.create_string_buffer(1 << 16)
    struct.pack_into("<I", buf, 0, tag)
    return buf


def call(lib, name, nparams, handle):
    """name(handle, 0, 0, ...): a fresh untyped function object whose arguments are all pointers
    (on x86-64 the callee reads only the handle before it returns)."""
    raw = C.CDLL(lib._name)[name]
    raw.argtypes = [C.c_void_p] * nparams
    raw.restype = C.c_void_p if name == "lbm_stream" else C.c_int
    return raw(None if handle is None else C.cast(handle, C.c_void_p), *([None] * (nparams - 1)))


def test_header_lists_the_handle_functions():
    ctx = exports_taking("lbm_ctx")
    ens = exports_taking("lbm_ens")
    assert len(ctx) >= 30 and len(ens) >= 9
    assert ("lbm_step", 3) in ctx and ("lbm_ensemble_step", 3) in ens


@pytest.mark.parametrize("tag", [ENS_TAG, 0, CTX_TAG ^ 1])
def test_context_exports_reject_foreign_handles(lib, tag):
    h = fake_handle(tag)
    for name, n in exports_taking("lbm_ctx"):
        if name == "lbm_last_error":
            continue
        rc = call(lib, name, n, h)
        if name == "lbm_stream":
            assert rc is None, name
        else:
            assert rc == LBM_EINVAL, f"{name} returned {rc}"
            assert b"not a live lbm_ctx handle" in lib.lbm_last_error(None), name
    # the message of a foreign handle is the thread's last message, never read from the handle
    assert b"lbm_ctx" in lib.lbm_last_error(C.cast(h, C.c_void_p))


@pytest.mark.parametrize("tag", [CTX_TAG, 0])
def test_ensemble_exports_reject_foreign_handles(lib, tag):
    h = fake_handle(tag)
    for name, n in exports_taking("lbm_ens"):
        if name == "lbm_ensemble_last_error":
            continue
        rc = call(lib, name, n, h)
        assert rc == LBM_EINVAL, f"{name} returned {rc}"
        assert b"not a live lbm_ens handle" in lib.lbm_last_error(None), name
    assert b"lbm_ens" in lib.lbm_ensemble_last_error(C.cast(h, C.c_void_p))


def test_null_handles(lib):
    for name, n in exports_taking("lbm_ctx"):
        if name in ("lbm_last_error", "lbm_destroy", "lbm_stream"):
            continue
        assert call(lib, name, n, None) == LBM_EINVAL, name
    for name, n in exports_taking("lbm_ens"):
        if name in ("lbm_ensemble_last_error", "lbm_ensemble_destroy"):
            continue
        assert call(lib, name, n, None) == LBM_EINVAL, name
    assert lib.lbm_destroy(None) == 0 and lib.lbm_ensemble_destroy(None) == 0


def test_ensemble_of_foreign_members_is_rejected(lib):
    h = fake_handle(ENS_TAG)
    arr = (C.c_void_p * 1)(C.cast(h, C.c_void_p).value)
    out = C.c_void_p()
    assert lib.lbm_ensemble_create(arr, 1, C.byref(out)) == LBM_EINVAL
    assert out.value is None
    assert b"member 0 is not a live lbm_ctx handle" in lib.lbm_last_error(None)
