"""Python front end of liblbm_hip.so (include/lbm_hip.h) through ctypes.

``Domain`` mirrors the reference driver's use of its per-step operators (src/main.F90):
    set_state        readrestart -> macrovars -> fB = fA      (src/main.F90:255-258)
    init_equilibrium fequil3, fillghosts, fA = fB, 2x boundarycond (src/main.F90:243-247)
    step             the time-loop body                        (src/main.F90:290-329)
    get_state        f1 as saverestart/testing write it        (src/m_saverestart.F90:125-136)
    macrovars        rho,u,v,w of f1 incl. ghosts              (src/m_macrovars.F90:4)
Arrays use the reference's layouts (C-contiguous numpy views of the Fortran arrays):
    f (nz+2, ny+2, nx+2, q), cell fields (nz+2, ny+2, nx+2), forcing (nz+2, ny+2, nx+2, 3),
    turbulence (nrturb+1, nz, ny).
The library must be built (python -m lbm_amd.build); there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbm_hip.so")

STORAGE = {"f32": 0, "f64": 1, "f16s": 2, "bf16s": 3, "f64t": 4}
BND_HALO = 3   # jbnd: ghost planes of j filled by halo_unpack (a domain tiled along j)
PART_EDGE, PART_INNER = 1, 2   # step_part: the cells next to the halos, then the others

EXPORTS = ["lbm_create", "lbm_destroy", "lbm_last_error", "lbm_set_state", "lbm_init_equilibrium",
           "lbm_get_state", "lbm_get_macros", "lbm_get_tau", "lbm_set_blanking", "lbm_set_ablvisc",
           "lbm_set_inflow", "lbm_set_inflow_series", "lbm_set_turbulence", "lbm_set_external_forcing",
           "lbm_step", "lbm_synchronize", "lbm_stream", "lbm_set_timing", "lbm_timing_read",
           "lbm_get_totals", "lbm_get_totals_noalloc", "lbm_get_info",
           "lbm_ensemble_create", "lbm_ensemble_destroy", "lbm_ensemble_last_error", "lbm_ensemble_step",
           "lbm_ensemble_synchronize", "lbm_ensemble_set_launch", "lbm_ensemble_get_launch",
           "lbm_ensemble_tune", "lbm_ensemble_set_timing", "lbm_ensemble_timing_read",
           "lbm_objective_run", "lbm_adjoint_run", "lbm_set_tangent", "lbm_objective_tangent",
           "lbm_generate_turbulence", "lbm_set_averaging", "lbm_get_averages",
           "lbm_halo_bytes", "lbm_halo_pack", "lbm_halo_unpack", "lbm_step_part",
           "lbm_set_turbines", "lbm_get_turbine_theta", "lbm_set_buoyancy"]


class Desc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("q", C.c_int32),
                ("storage", C.c_int32), ("ibgk", C.c_int32), ("ihrr", C.c_int32), ("ivreman", C.c_int32),
                ("ibnd", C.c_int32), ("jbnd", C.c_int32), ("kbnd", C.c_int32), ("iturb_pos", C.c_int32),
                ("store_tau", C.c_int32), ("block_x", C.c_int32), ("block_y", C.c_int32),
                ("kinevisc", C.c_double), ("smagorinsky", C.c_double), ("rho0", C.c_double)]


class Turbines(C.Structure):
    """lbm_turbines (include/lbm_hip.h)."""
    _fields_ = [("nturbines", C.c_int32), ("nblades", C.c_int32), ("nchords", C.c_int32), ("nfoils", C.c_int32),
                ("nang_max", C.c_int32), ("reserved", C.c_int32)] + \
               [(n, C.c_void_p) for n in ("hub", "yaw", "tilt", "omegand", "pitch", "theta", "relm", "dc", "chord",
                                          "twist", "nfoil", "nang", "foil")]


class LBMError(RuntimeError):
    pass


_lib = None


def load_library(path: str = LIB_PATH):
    """Load liblbm_hip.so and declare its signatures.  Raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise LBMError(f"{path} not found: build it with `python -m lbm_amd.build`")
    lib = C.CDLL(path)
    vp, i32, dbl = C.c_void_p, C.c_int, C.c_double
    sig = {
        "lbm_create": ([C.POINTER(Desc), C.POINTER(vp)], i32),
        "lbm_destroy": ([vp], i32),
        "lbm_last_error": ([vp], C.c_char_p),
        "lbm_set_state": ([vp, vp, i32], i32),
        "lbm_init_equilibrium": ([vp, vp, vp, vp, vp, i32, i32], i32),
        "lbm_get_state": ([vp, vp, i32], i32),
        "lbm_get_macros": ([vp, vp, vp, vp, vp, i32], i32),
        "lbm_get_tau": ([vp, vp, i32], i32),
        "lbm_set_blanking": ([vp, vp], i32),
        "lbm_set_ablvisc": ([vp, vp, i32], i32),
        "lbm_set_inflow": ([vp, vp, dbl, dbl, i32], i32),
        "lbm_set_inflow_series": ([vp, i32, i32, vp, vp, i32], i32),
        "lbm_set_turbulence": ([vp, vp, vp, vp, i32, dbl, i32], i32),
        "lbm_set_external_forcing": ([vp, vp, i32], i32),
        "lbm_step": ([vp, i32, i32], i32),
        "lbm_synchronize": ([vp], i32),
        "lbm_stream": ([vp], vp),
        "lbm_halo_bytes": ([vp, C.POINTER(C.c_size_t)], i32),
        "lbm_halo_pack": ([vp, i32, vp], i32),
        "lbm_halo_unpack": ([vp, i32, vp], i32),
        "lbm_step_part": ([vp, i32, i32], i32),
        "lbm_set_timing": ([vp, i32], i32),
        "lbm_timing_read": ([vp, C.POINTER(dbl), C.POINTER(i32), C.POINTER(dbl), C.POINTER(i32)], i32),
        "lbm_get_totals": ([vp, C.POINTER(dbl), C.POINTER(dbl)], i32),
        "lbm_set_buoyancy": ([vp, vp, dbl, dbl, i32], i32),
        "lbm_get_totals_noalloc": ([vp, C.POINTER(dbl), C.POINTER(dbl)], i32),
        "lbm_get_info": ([vp, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_int64)], i32),
        "lbm_ensemble_create": ([C.POINTER(vp), i32, C.POINTER(vp)], i32),
        "lbm_ensemble_destroy": ([vp], i32),
        "lbm_ensemble_last_error": ([vp], C.c_char_p),
        "lbm_ensemble_step": ([vp, i32, i32], i32),
        "lbm_ensemble_synchronize": ([vp], i32),
        "lbm_ensemble_set_launch": ([vp, i32, i32], i32),
        "lbm_ensemble_get_launch": ([vp, C.POINTER(i32), C.POINTER(i32)], i32),
        "lbm_ensemble_tune": ([vp, i32, i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(dbl)], i32),
        "lbm_ensemble_set_timing": ([vp, i32], i32),
        "lbm_ensemble_timing_read": ([vp, C.POINTER(dbl), C.POINTER(i32)], i32),
        "lbm_objective_run": ([vp, i32, i32, C.POINTER(C.c_int32), C.POINTER(dbl)], i32),
        "lbm_adjoint_run": ([vp, i32, i32, i32, C.POINTER(C.c_int32), C.POINTER(dbl), vp, vp], i32),
        "lbm_set_tangent": ([vp, vp, vp], i32),
        "lbm_set_averaging": ([vp, i32], i32),
        "lbm_get_averages": ([vp, vp, vp, vp, vp, C.POINTER(i32), i32], i32),
        "lbm_generate_turbulence": ([vp, i32, dbl, C.c_uint64, vp, vp, vp, vp, i32], i32),
        "lbm_objective_tangent": ([vp, C.POINTER(dbl)], i32),
        "lbm_set_turbines": ([vp, C.POINTER(Turbines)], i32),
        "lbm_get_turbine_theta": ([vp, C.POINTER(dbl)], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def buoyancy_scaling(length=1.0, vel=1.0, dtype=np.float32):
    """scaling = g * p2l%time**2 / p2l%length with g = 9.81 and p2l%time = length / vel, every
    operation rounded in dtype as the reference computes it (src/m_buoyancy_forcing.F90:33-34,
    src/m_readinfile.F90:171)."""
    g, L, V = dtype(9.81), dtype(length), dtype(vel)
    t = dtype(L / V)
    return float(dtype(dtype(g * dtype(t * t)) / L))


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Domain:
    """One lattice-update context on the current HIP device."""

    def __init__(self, q, dims, storage="f32", ibgk=2, ihrr=0, ivreman=0, smagorinsky=0.0,
                 kinevisc=1.0 / 30.0, ibnd=0, jbnd=0, kbnd=0, rho0=1.0, iturb_pos=10, store_tau=False,
                 block=(0, 0)):
        self.lib = load_library()
        nx, ny, nz = dims
        self.q, self.dims, self.storage = q, (nx, ny, nz), storage
        self.real = np.float64 if storage in ("f64", "f64t") else np.float32
        d = Desc(nx, ny, nz, q, STORAGE[storage], ibgk, ihrr, ivreman, ibnd, jbnd, kbnd, iturb_pos,
                 int(store_tau), block[0], block[1], float(kinevisc), float(smagorinsky), float(rho0))
        h = C.c_void_p()
        rc = self.lib.lbm_create(C.byref(d), C.byref(h))
        if rc != 0:
            raise LBMError(f"lbm_create failed ({rc}): {self.lib.lbm_last_error(None).decode()}")
        self.h = h
        self.desc = d

    # -- plumbing -----------------------------------------------------------------------
    def _chk(self, rc, what):
        if rc != 0:
            raise LBMError(f"{what} failed ({rc}): {self.lib.lbm_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.lbm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def shape_f(self):
        nx, ny, nz = self.dims
        return (nz + 2, ny + 2, nx + 2, self.q)

    @property
    def shape_cell(self):
        nx, ny, nz = self.dims
        return (nz + 2, ny + 2, nx + 2)

    def _arr(self, a, shape, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype or (a.dtype if a.dtype in (np.float32, np.float64) else np.float64))
        if tuple(a.shape) != tuple(shape):
            raise ValueError(f"expected shape {shape}, got {a.shape}")
        return a

    # -- state --------------------------------------------------------------------------
    def set_state(self, f):
        f = self._arr(f, self.shape_f)
        self._chk(self.lib.lbm_set_state(self.h, _p(f), f.itemsize), "lbm_set_state")

    def init_equilibrium(self, rho, u, v, w, it0=0):
        arrs = [self._arr(a, self.shape_cell, np.float64) for a in (rho, u, v, w)]
        self._chk(self.lib.lbm_init_equilibrium(self.h, *[_p(a) for a in arrs], 8, it0), "lbm_init_equilibrium")

    def get_state(self, dtype=None):
        out = np.empty(self.shape_f, dtype=dtype or self.real)
        self._chk(self.lib.lbm_get_state(self.h, _p(out), out.itemsize), "lbm_get_state")
        return out

    def macrovars(self, dtype=None):
        dt = dtype or self.real
        r, u, v, w = (np.empty(self.shape_cell, dt) for _ in range(4))
        self._chk(self.lib.lbm_get_macros(self.h, _p(r), _p(u), _p(v), _p(w), r.itemsize), "lbm_get_macros")
        return r, u, v, w

    def tau(self):
        t = np.empty(self.shape_cell, self.real)
        self._chk(self.lib.lbm_get_tau(self.h, _p(t), t.itemsize), "lbm_get_tau")
        return t

    def set_averaging(self, on=True):
        """Time averaging of the moments after each following step (lbm_set_averaging)."""
        self._chk(self.lib.lbm_set_averaging(self.h, int(on)), "lbm_set_averaging")

    def averages(self, dtype=None):
        """(uave, vave, wave, Ti, count) with cell-field shape (nz+2, ny+2, nx+2) (lbm_get_averages)."""
        dt = dtype or self.real
        out = [np.empty(self.shape_cell, dt) for _ in range(4)]
        n = C.c_int()
        self._chk(self.lib.lbm_get_averages(self.h, *[_p(a) for a in out], C.byref(n), np.dtype(dt).itemsize),
                  "lbm_get_averages")
        return (*out, n.value)

    def totals(self, noalloc=False):
        """Mass and momentum (FP64); noalloc: through lbm_get_totals_noalloc (no device allocation)."""
        m = C.c_double()
        mom = (C.c_double * 3)()
        name = "lbm_get_totals_noalloc" if noalloc else "lbm_get_totals"
        self._chk(getattr(self.lib, name)(self.h, C.byref(m), mom), name)
        return m.value, np.array(mom[:])

    def info(self):
        b, p, pl = C.c_int64(), C.c_int32(), C.c_int64()
        self._chk(self.lib.lbm_get_info(self.h, C.byref(b), C.byref(p), C.byref(pl)), "lbm_get_info")
        return dict(device_bytes=b.value, row_pitch=p.value, plane=pl.value)

    # -- inputs -------------------------------------------------------------------------
    def set_blanking(self, mask):
        if mask is None:
            self._chk(self.lib.lbm_set_blanking(self.h, None), "lbm_set_blanking")
            return
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        if m.shape != self.shape_cell:
            raise ValueError("mask shape")
        self._chk(self.lib.lbm_set_blanking(self.h, _p(m)), "lbm_set_blanking")

    def set_ablvisc(self, abl):
        a = np.ascontiguousarray(abl, dtype=np.float64)
        self._chk(self.lib.lbm_set_ablvisc(self.h, _p(a), 8), "lbm_set_ablvisc")

    def set_inflow(self, uini, udir=0.0, shear=None):
        s = None if shear is None else np.ascontiguousarray(shear, dtype=np.float64)
        self._chk(self.lib.lbm_set_inflow(self.h, _p(s), float(uini), float(udir), 8), "lbm_set_inflow")

    def set_inflow_series(self, it_lo, uvel_time, udir_time):
        ut = np.ascontiguousarray(uvel_time, dtype=np.float64)
        ud = np.ascontiguousarray(udir_time, dtype=np.float64)
        self._chk(self.lib.lbm_set_inflow_series(self.h, it_lo, it_lo + len(ut) - 1, _p(ut), _p(ud), 8),
                  "lbm_set_inflow_series")

    def set_turbulence(self, uu, vv, ww, nrturb, ampl):
        if uu is None:
            self._chk(self.lib.lbm_set_turbulence(self.h, None, None, None, 0, 0.0, 8), "lbm_set_turbulence")
            return
        dt = np.float64 if self.real == np.float64 else np.float32
        a = [np.ascontiguousarray(x, dtype=dt) for x in (uu, vv, ww)]
        self._chk(self.lib.lbm_set_turbulence(self.h, *[_p(x) for x in a], int(nrturb), float(ampl),
                                              a[0].itemsize), "lbm_set_turbulence")

    def generate_turbulence(self, nrturb, ampl, seed, return_fields=True):
        """Inflow-turbulence fields generated on the device (lbm_generate_turbulence) and made
        the forcing; returns (uu, vv, ww, rr), each (nrturb+1, nz, ny), when return_fields."""
        nx, ny, nz = self.dims
        out = [np.empty((nrturb + 1, nz, ny), self.real) for _ in range(4)] if return_fields else [None] * 4
        self._chk(self.lib.lbm_generate_turbulence(self.h, nrturb, float(ampl), int(seed) & (2 ** 64 - 1),
                                                   *[_p(a) for a in out], np.dtype(self.real).itemsize),
                  "lbm_generate_turbulence")
        return tuple(out) if return_fields else None

    def set_external_forcing(self, force):
        if force is None:
            self._chk(self.lib.lbm_set_external_forcing(self.h, None, 8), "lbm_set_external_forcing")
            return
        F = self._arr(force, self.shape_cell + (3,))
        self._chk(self.lib.lbm_set_external_forcing(self.h, _p(F), F.itemsize), "lbm_set_external_forcing")

    def set_buoyancy(self, pottemp, theta0=300.0, scaling=None, length=1.0, vel=1.0):
        """Buoyancy forcing from the potential temperature field (cell layout); None switches it
        off.  scaling defaults to the reference's g * p2l%time^2 / p2l%length (g = 9.81,
        p2l%time = length / vel) in the compute type's rounding."""
        if pottemp is None:
            self._chk(self.lib.lbm_set_buoyancy(self.h, None, 300.0, 0.0, 8), "lbm_set_buoyancy")
            return
        if scaling is None:
            scaling = buoyancy_scaling(length, vel, np.float64 if self.real == np.float64 else np.float32)
        T = self._arr(pottemp, self.shape_cell)
        self._chk(self.lib.lbm_set_buoyancy(self.h, _p(T), float(theta0), float(scaling), T.itemsize),
                  "lbm_set_buoyancy")

    def set_turbines(self, t):
        """Actuator-line turbines (lbm_set_turbines) from the dict of lbm_amd.turbines.from_infile;
        None removes them."""
        if t is None:
            self._chk(self.lib.lbm_set_turbines(self.h, None), "lbm_set_turbines")
            return
        f64 = lambda a: np.ascontiguousarray(a, np.float64)
        foils = t["foils"]
        nmax = max(len(f) for f in foils)
        foil = np.zeros((len(foils), nmax, 3))
        for k, f in enumerate(foils):
            foil[k, :len(f)] = f
        arr = dict(hub=f64(t["hub"]).reshape(-1), yaw=f64(t["yaw"]), tilt=f64(t["tilt"]), omegand=f64(t["omegand"]),
                   pitch=f64(t["pitch"]), theta=f64(t["theta"]), relm=f64(t["relm"]), dc=f64(t["dc"]),
                   chord=f64(t["chord"]), twist=f64(t["twist"]), nfoil=np.ascontiguousarray(t["nfoil"], np.int32),
                   nang=np.array([len(f) for f in foils], np.int32), foil=foil)
        d = Turbines(len(arr["yaw"]), int(t.get("nblades", 3)), len(arr["relm"]), len(foils), nmax, 0,
                     *[arr[k].ctypes.data for k in ("hub", "yaw", "tilt", "omegand", "pitch", "theta", "relm", "dc",
                                                     "chord", "twist", "nfoil", "nang", "foil")])
        self._chk(self.lib.lbm_set_turbines(self.h, C.byref(d)), "lbm_set_turbines")
        self._nturb = len(arr["yaw"])

    def turbine_theta(self):
        """The rotors' azimuths [rad] after the last step."""
        th = (C.c_double * self._nturb)()
        self._chk(self.lib.lbm_get_turbine_theta(self.h, th), "lbm_get_turbine_theta")
        return np.array(th[:])

    # -- objective and adjoint -------------------------------------------------------------
    @staticmethod
    def _box(box):
        (i0, i1), (j0, j1), (k0, k1) = box
        return (C.c_int32 * 6)(i0, i1, j0, j1, k0, k1)

    def objective(self, it0, nsteps, box):
        """Steps it0+1 .. it0+nsteps; returns the time-mean kinetic energy 0.5 rho|u|^2 over box
        ((i0,i1),(j0,j1),(k0,k1)), 1-based inclusive (lbm_objective_run)."""
        J = C.c_double()
        self._chk(self.lib.lbm_objective_run(self.h, it0, nsteps, self._box(box), C.byref(J)), "lbm_objective_run")
        return J.value

    def adjoint(self, it0, nsteps, every, box, grad_force=True):
        """As objective(), plus the reverse sweep: returns (J, dJ/d external forcing with the
        forcing's shape (nz+2, ny+2, nx+2, 3) or None, dJ/d uvel_shear (nz,))."""
        nx, ny, nz = self.dims
        J = C.c_double()
        gF = np.zeros((nz + 2, ny + 2, nx + 2, 3)) if grad_force else None
        gS = np.zeros(nz)
        self._chk(self.lib.lbm_adjoint_run(self.h, it0, nsteps, every, self._box(box), C.byref(J), _p(gF), _p(gS)),
                  "lbm_adjoint_run")
        return J.value, gF, gS

    def set_tangent(self, dforce=None, dshear=None):
        """Direction of the tangent run (storage "f64t"): d forcing (forcing's shape) and d shear."""
        F = None if dforce is None else self._arr(dforce, self.shape_cell + (3,), np.float64)
        S = None if dshear is None else np.ascontiguousarray(dshear, dtype=np.float64)
        self._chk(self.lib.lbm_set_tangent(self.h, _p(F), _p(S)), "lbm_set_tangent")

    def objective_tangent(self):
        dJ = C.c_double()
        self._chk(self.lib.lbm_objective_tangent(self.h, C.byref(dJ)), "lbm_objective_tangent")
        return dJ.value

    # -- time stepping -------------------------------------------------------------------
    def step(self, it0, nsteps=1):
        self._chk(self.lib.lbm_step(self.h, int(it0), int(nsteps)), "lbm_step")

    def step_part(self, it, part):
        """Step ``it`` of a halo domain in two launches: PART_EDGE (j = 1, 2, ny-1, ny; j = 1, ny
        without a closed face), after which halo_pack may run, then PART_INNER (the other cells),
        which completes the step."""
        self._chk(self.lib.lbm_step_part(self.h, int(it), int(part)), "lbm_step_part")

    def synchronize(self):
        self._chk(self.lib.lbm_synchronize(self.h), "lbm_synchronize")

    def stream_ptr(self):
        """The context's hipStream_t (as an integer), e.g. for torch.cuda.ExternalStream."""
        return self.lib.lbm_stream(self.h) or 0

    def halo_bytes(self):
        """Size of one halo buffer (jbnd = BND_HALO domains)."""
        n = C.c_size_t()
        self._chk(self.lib.lbm_halo_bytes(self.h, C.byref(n)), "lbm_halo_bytes")
        return n.value

    def halo_pack(self, side, dev_ptr):
        """Enqueue the pack of plane j = 1 (side 0) or j = ny (side 1) into device memory dev_ptr."""
        self._chk(self.lib.lbm_halo_pack(self.h, int(side), C.c_void_p(dev_ptr)), "lbm_halo_pack")

    def halo_unpack(self, side, dev_ptr):
        """Enqueue the unpack of a neighbour's halo into ghost plane j = 0 (side 0) or ny+1 (side 1)."""
        self._chk(self.lib.lbm_halo_unpack(self.h, int(side), C.c_void_p(dev_ptr)), "lbm_halo_unpack")

    def set_timing(self, on=True, every=1):
        """HIP-event timing of the step kernel (and boundary kernels) of every ``every``-th step."""
        self._chk(self.lib.lbm_set_timing(self.h, int(every) if on else 0), "lbm_set_timing")

    def timing_read(self):
        s, b = C.c_double(), C.c_double()
        ns, nb = C.c_int(), C.c_int()
        self._chk(self.lib.lbm_timing_read(self.h, C.byref(s), C.byref(ns), C.byref(b), C.byref(nb)),
                  "lbm_timing_read")
        return dict(step_ms=s.value, n_step=ns.value, bc_ms=b.value, n_bc=nb.value)


class Ensemble:
    """Members (Domain objects of one lattice, storage and collision model) advanced together by
    the batched kernels (lbm_ensemble_* in include/lbm_hip.h).  The members stay usable as
    Domains between ensemble steps and must outlive the ensemble."""

    def __init__(self, members):
        self.lib = load_library()
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m.h.value for m in self.members])
        h = C.c_void_p()
        rc = self.lib.lbm_ensemble_create(arr, len(self.members), C.byref(h))
        if rc != 0:
            msg = self.lib.lbm_ensemble_last_error(h).decode()   # h NULL: the thread's last message
            if h.value:
                self.lib.lbm_ensemble_destroy(h)
            raise LBMError(f"lbm_ensemble_create failed ({rc}): {msg}")
        self.h = h

    def _chk(self, rc, what):
        if rc != 0:
            raise LBMError(f"{what} failed ({rc}): {self.lib.lbm_ensemble_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.lbm_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- time stepping -------------------------------------------------------------------
    def step(self, it0, nsteps=1):
        self._chk(self.lib.lbm_ensemble_step(self.h, it0, nsteps), "lbm_ensemble_step")

    def synchronize(self):
        self._chk(self.lib.lbm_ensemble_synchronize(self.h), "lbm_ensemble_synchronize")

    def set_launch(self, block, blocks_per_cu):
        self._chk(self.lib.lbm_ensemble_set_launch(self.h, block, blocks_per_cu), "lbm_ensemble_set_launch")

    def launch(self):
        b, p = C.c_int(), C.c_int()
        self._chk(self.lib.lbm_ensemble_get_launch(self.h, C.byref(b), C.byref(p)), "lbm_ensemble_get_launch")
        return b.value, p.value

    def tune(self, it0, trial_steps=10):
        """Returns (block, blocks_per_cu, ms_per_step) of the fastest launch shape; the members'
        states are unchanged."""
        b, p, ms = C.c_int(), C.c_int(), C.c_double()
        self._chk(self.lib.lbm_ensemble_tune(self.h, it0, trial_steps, C.byref(b), C.byref(p), C.byref(ms)),
                  "lbm_ensemble_tune")
        return b.value, p.value, ms.value

    def set_timing(self, on=True):
        self._chk(self.lib.lbm_ensemble_set_timing(self.h, int(on)), "lbm_ensemble_set_timing")

    def timing_read(self):
        ms, n = C.c_double(), C.c_int()
        self._chk(self.lib.lbm_ensemble_timing_read(self.h, C.byref(ms), C.byref(n)), "lbm_ensemble_timing_read")
        return ms.value, n.value
