"""Fortran unformatted sequential records (lbm_amd.fortran_io) in both conventions for payloads of
2 GiB and more (src/m_saverestart.F90:125-136 writes the state as one record):
* gfortran splits a payload into sub-records of at most 2^31 - 9 bytes with negative continuation
  markers: exercised here with the sub-record limit lowered, so a small state needs a chain;
* flang/amdflang writes one record whose markers hold the length modulo 2^32: a sparse file with a
  payload just above 4 GiB, read back with the expected length.
Neither needs a device."""
import os
import struct

import numpy as np
import pytest

from lbm_amd import fortran_io as fio


def small_state(dtype=np.float32):
    rng = np.random.default_rng(0)
    return rng.random((5, 6, 7, 19)).astype(dtype)    # (nz+2, ny+2, nx+2, nl)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_gfortran_subrecord_chain_round_trip(tmp_path, monkeypatch, dtype):
    f = small_state(dtype)
    payload = 16 + f.nbytes
    limit = 1000                                       # bytes per sub-record (gfortran: 2^31 - 9)
    monkeypatch.setattr(fio, "_GF_MAX", limit)
    p = str(tmp_path / "restart.uf")
    fio.write_restart(p, f, subrecords=True)
    # the chain's structure: ceil(payload/limit) sub-records; leading markers negative except the
    # last, trailing markers negative except the first's
    raw = open(p, "rb").read()
    n_sub = -(-payload // limit)
    assert len(raw) == payload + 8 * n_sub
    off, leads, trails = 0, [], []
    while off < len(raw):
        (m,) = struct.unpack_from("<i", raw, off)
        n = abs(m)
        (t,) = struct.unpack_from("<i", raw, off + 4 + n)
        leads.append(m)
        trails.append(t)
        off += 8 + n
    assert len(leads) == n_sub
    assert all(m < 0 for m in leads[:-1]) and leads[-1] > 0
    assert trails[0] > 0 and all(t < 0 for t in trails[1:])
    g = fio.read_restart(p)
    assert g.dtype == f.dtype and np.array_equal(g, f)


def test_flang_record_above_4_gib(tmp_path):
    """One record of 4 GiB + 64 bytes: both markers read 64 (the length modulo 2^32); the reader
    takes the length it expects and returns the whole payload (a sparse file: only its first and
    last bytes are stored)."""
    psutil = pytest.importorskip("psutil")
    if psutil.virtual_memory().available < 12 * 2**30:
        pytest.skip("needs about 5 GiB of free memory")
    n = 2**32 + 64
    p = str(tmp_path / "big.uf")
    with open(p, "wb") as fh:
        fh.write(struct.pack("<I", n % 2**32))
        fh.write(b"HEAD")
        fh.seek(4 + n - 4)
        fh.write(b"TAIL")
        fh.write(struct.pack("<I", n % 2**32))
    assert os.path.getsize(p) == n + 8
    with open(p, "rb") as fh:
        with pytest.raises(ValueError):
            fio.read_record(fh, expected=n - 8)        # a length that the markers contradict
    with open(p, "rb") as fh:
        data = fio.read_record(fh, expected=n)
        assert fh.read() == b""                        # positioned after the trailing marker
    assert len(data) == n and data[:4] == b"HEAD" and data[-4:] == b"TAIL"
    del data


def test_restart_reader_picks_the_precision_from_the_record_length(tmp_path):
    for dt in (np.float32, np.float64):
        f = small_state(dt)
        p = str(tmp_path / f"r_{np.dtype(dt).itemsize}.uf")
        fio.write_restart(p, f)
        g = fio.read_restart(p)
        assert g.dtype == dt and np.array_equal(g, f)


def test_theta_and_turbulence_records(tmp_path):
    th = np.array([0.25, 1.5, -3.0], np.float32)
    fio.write_theta(str(tmp_path / "theta.uf"), th)
    assert np.array_equal(fio.read_theta(str(tmp_path / "theta.uf")), th)
    rng = np.random.default_rng(1)
    uu, vv, ww, rr = (rng.random((4, 3, 5)).astype(np.float32) for _ in range(4))   # (nrturb+1, nz, ny)
    fio.write_turbulence(str(tmp_path / "t.uf"), uu, vv, ww, rr)
    got = fio.read_turbulence(str(tmp_path / "t.uf"))
    assert all(np.array_equal(a, b) for a, b in zip(got, (uu, vv, ww, rr)))
