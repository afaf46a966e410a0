"""Fortran unformatted sequential records, as the reference's restart files use them.

The reference writes its state with ``write(iunit) nx,ny,nz,nl,f`` on a file opened
``form="unformatted"`` (src/m_saverestart.F90:125-136) and reads it back the same way
(src/m_readrestart.F90:163-185).  A sequential unformatted record is

    [int32 marker][payload bytes][int32 marker]

Two conventions exist for payloads of 2 GiB or more:

* flang/amdflang (the reference's CPU build here): one record whose markers hold the
  payload length modulo 2**32;
* gfortran: the payload is split into sub-records of at most 2**31-9 bytes; every
  sub-record except the last has a negative leading marker (and the trailing marker of
  every sub-record except the first is negative).

``read_records`` accepts both; ``write_record`` writes the flang form by default and the
gfortran form with ``subrecords=True``.
"""
from __future__ import annotations

import os
import struct

import numpy as np

_GF_MAX = 2**31 - 9


def write_record(fh, payload: bytes | memoryview, subrecords: bool = False) -> None:
    n = len(payload)
    if not subrecords:
        m = struct.pack("<I", n % 2**32)
        fh.write(m)
        fh.write(payload)
        fh.write(m)
        return
    mv = memoryview(payload)
    off = 0
    first = True
    while True:
        chunk = min(_GF_MAX, n - off)
        last = off + chunk >= n
        lead = chunk if last else -chunk
        trail = chunk if first else -chunk
        fh.write(struct.pack("<i", lead))
        fh.write(mv[off:off + chunk])
        fh.write(struct.pack("<i", trail))
        off += chunk
        first = False
        if last:
            break


def read_record(fh, expected: int | None = None) -> bytes:
    """Read one record.  ``expected`` (payload bytes) resolves flang records >= 4 GiB."""
    head = fh.read(4)
    if len(head) < 4:
        raise EOFError("end of file")
    (m,) = struct.unpack("<i", head)
    if m < 0 and expected is None:
        # gfortran continuation chain
        parts = []
        while True:
            n = -m if m < 0 else m
            parts.append(fh.read(n))
            fh.read(4)
            if m >= 0:
                break
            (m,) = struct.unpack("<i", fh.read(4))
        return b"".join(parts)
    um = m % 2**32
    if expected is not None and um != expected % 2**32:
        # could still be a gfortran chain for a large record
        if m < 0:
            fh.seek(-4, os.SEEK_CUR)
            return read_record(fh, None)
        raise ValueError(f"record marker {um} does not match expected payload {expected}")
    n = expected if expected is not None else um
    data = fh.read(n)
    if len(data) != n:
        raise ValueError("truncated record")
    fh.read(4)
    return data


def read_records(path: str):
    out = []
    with open(path, "rb") as fh:
        while True:
            try:
                out.append(read_record(fh))
            except EOFError:
                break
    return out


def _peek_ints(fh, count: int):
    pos = fh.tell()
    fh.read(4)
    vals = struct.unpack("<" + "i" * count, fh.read(4 * count))
    fh.seek(pos)
    return vals


# ---------------------------------------------------------------------------------------
# restart/restart_<tile4>_<it6>.uf : [nx,ny,nz,nl][f(nl,0:nx+1,0:ny+1,0:nz+1)]
# ---------------------------------------------------------------------------------------
def restart_name(it: int, tile: int = 0, prefix: str = "restart") -> str:
    """File name the reference uses (src/m_saverestart.F90:57-60,126-127)."""
    return f"{prefix}_{tile:04d}_{it:06d}.uf"


def write_restart(path: str, f: np.ndarray, subrecords: bool = False) -> None:
    """``f`` has the reference's shape (nl, nx+2, ny+2, nz+2) in Fortran order, i.e. a
    C-contiguous numpy array of shape (nz+2, ny+2, nx+2, nl)."""
    nzp, nyp, nxp, nl = f.shape
    hdr = struct.pack("<4i", nxp - 2, nyp - 2, nzp - 2, nl)
    body = np.ascontiguousarray(f).tobytes()
    with open(path, "wb") as fh:
        write_record(fh, hdr + body, subrecords)


def read_restart(path: str) -> np.ndarray:
    """Returns f as a C-contiguous (nz+2, ny+2, nx+2, nl) array of float32 or float64,
    chosen from the record length."""
    with open(path, "rb") as fh:
        nx, ny, nz, nl = _peek_ints(fh, 4)
        ncell = nl * (nx + 2) * (ny + 2) * (nz + 2)
        size = os.path.getsize(path)
        # the file holds one record (8 bytes of markers: flang at any size, gfortran below 2 GiB) or
        # a gfortran chain of ceil(payload / (2^31 - 9)) sub-records (8 bytes of markers each)
        for dt in (np.float32, np.float64):
            expected = 16 + ncell * np.dtype(dt).itemsize
            if size == expected + 8:
                chain = False
                break
            if size == expected + 8 * (-(-expected // _GF_MAX)):
                chain = True
                break
        else:
            raise ValueError(f"{path}: size {size} fits neither FP32 nor FP64 for {nx}x{ny}x{nz}x{nl}")
        data = read_record(fh, None if chain else expected)
    return np.frombuffer(data[16:], dtype=dt).reshape(nz + 2, ny + 2, nx + 2, nl).copy()


def write_turbulence(path: str, uu: np.ndarray, vv: np.ndarray, ww: np.ndarray, rr: np.ndarray,
                     subrecords: bool = False) -> None:
    """restart/turbulence_<tile>_<it>.uf: [ny,nz,nrturb][uu,vv,ww,rr (ny,nz,0:nrturb)]
    (src/m_saverestart.F90:66-83).  Arrays are C-contiguous (nrturb+1, nz, ny)."""
    nt1, nz, ny = uu.shape
    hdr = struct.pack("<3i", ny, nz, nt1 - 1)
    body = b"".join(np.ascontiguousarray(a).tobytes() for a in (uu, vv, ww, rr))
    with open(path, "wb") as fh:
        write_record(fh, hdr + body, subrecords)


def read_turbulence(path: str, dtype=np.float32):
    with open(path, "rb") as fh:
        ny, nz, nr = _peek_ints(fh, 3)
        n = ny * nz * (nr + 1)
        data = read_record(fh, 12 + 4 * n * np.dtype(dtype).itemsize)
    arr = np.frombuffer(data[12:], dtype=dtype).reshape(4, nr + 1, nz, ny)
    return tuple(a.copy() for a in arr)


def read_testing(path: str, shape) -> np.ndarray:
    """testing/testing_<tile>_<it>.uf holds real(kind=4) f including ghosts, no header
    (src/m_testing.F90:76-80).  ``shape`` = (nz+2, ny+2, nx+2, nl)."""
    with open(path, "rb") as fh:
        data = read_record(fh, 4 * int(np.prod(shape)))
    return np.frombuffer(data, dtype=np.float32).reshape(shape).copy()


# ---------------------------------------------------------------------------------------
# uvel_shear.dat / uvel_time.dat : the inflow profile and its time series
# ---------------------------------------------------------------------------------------
def _list_directed_rows(path: str):
    """Rows of whitespace- or comma-separated values, blank lines skipped (list-directed input)."""
    with open(path) as fh:
        return [r.replace(",", " ").split() for r in fh if r.strip()]


def read_inflow(directory: str, nz: int, nt0: int, nt1: int, uini: float, udir: float, vel: float = 1.0,
                time: float = 1.0, dtype=np.float32):
    """The inflow profile and time series the reference reads at start-up (src/m_inflow.F90:27-86),
    in the REAL kind ``dtype``: returns (uvel_shear (nz,), uvel_time (nt1-nt0+1,), udir_time (nt1-nt0+1,)),
    the series for steps nt0..nt1, ready for Domain.set_inflow / Domain.set_inflow_series.

    * uvel_shear.dat (optional): nz rows "k z u"; uvel_shear(k) = u(k) / u(nz), else 1.
    * uvel_time.dat (optional): rows "t u dir"; u is divided by ``vel`` (p2l%vel, m/s per lattice
      unit); step i is at t = (i - 1) * ``time`` (p2l%time); the series is u, dir of the first row
      up to its t, of the last row from its t on, linear in between.  Without the file every step
      takes uini and udir.
    All arithmetic rounds in ``dtype`` as the reference's REAL does."""
    R = np.dtype(dtype).type
    uh = np.ones(nz, dtype)
    p = os.path.join(directory, "uvel_shear.dat")
    if os.path.exists(p):
        rows = _list_directed_rows(p)
        if len(rows) < nz:
            raise ValueError(f"{p}: {len(rows)} rows for nz = {nz}")
        for k in range(nz):
            uh[k] = R(rows[k][2])
    shear = (uh / uh[nz - 1]).astype(dtype)
    n = nt1 - nt0 + 1
    ut = np.full(n, R(uini), dtype)
    ud = np.full(n, R(udir), dtype)
    p = os.path.join(directory, "uvel_time.dat")
    if os.path.exists(p):
        rows = _list_directed_rows(p)
        td = np.array([R(r[0]) for r in rows], dtype)
        uv = np.array([R(r[1]) for r in rows], dtype) / R(vel)
        dv = np.array([R(r[2]) for r in rows], dtype)
        for q, i in enumerate(range(nt0, nt1 + 1)):
            t = R(i - 1) * R(time)
            if t <= td[0]:
                ut[q], ud[q] = uv[0], dv[0]
            elif t >= td[-1]:
                ut[q], ud[q] = uv[-1], dv[-1]
            else:
                for m in range(len(td) - 1):
                    if td[m] <= t <= td[m + 1]:
                        s = (t - td[m]) / (td[m + 1] - td[m])
                        ut[q] = uv[m] + s * (uv[m + 1] - uv[m])
                        ud[q] = dv[m] + s * (dv[m + 1] - dv[m])
                        break
    return shear, ut, ud


def write_theta(path: str, theta, dtype=np.float32, subrecords: bool = False) -> None:
    """restart/theta_<tile>_<it>.uf: the rotor azimuths theta(nturbines) [rad], one record
    (src/m_saverestart.F90:118-122, read back by src/m_readrestart.F90:103-116)."""
    with open(path, "wb") as fh:
        write_record(fh, np.ascontiguousarray(theta, dtype=dtype).tobytes(), subrecords)


def read_theta(path: str, dtype=np.float32) -> np.ndarray:
    with open(path, "rb") as fh:
        return np.frombuffer(read_record(fh), dtype=dtype).copy()


def write_pottemp(path: str, pottemp, subrecords: bool = False) -> None:
    """restart/pottemp_<tile>_<it>.uf: the potential temperature pottemp(0:nx+1,0:ny+1,0:nz+1), one
    record (src/m_saverestart.F90:102-114, read back by src/m_readrestart.F90:119-136); the array is
    given C-contiguous as (nz+2, ny+2, nx+2) and keeps its dtype."""
    with open(path, "wb") as fh:
        write_record(fh, np.ascontiguousarray(pottemp).tobytes(), subrecords)


def read_pottemp(path: str, shape, dtype=np.float32) -> np.ndarray:
    with open(path, "rb") as fh:
        return np.frombuffer(read_record(fh), dtype=dtype).reshape(shape).copy()
