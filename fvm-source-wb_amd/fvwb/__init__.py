"""Python host mirror of the MI355X well-balanced Euler-with-gravity solver.

Thin ctypes layer over the C ABI in include/fvwb.h (libfvwb.so, built in-tree by
``make -C fvm-source-wb_amd/csrc``).  The names follow the reference's external
procedures (benchmark_2d.f90): ``compute_update_exact(u, w_eq)`` returns dudt,
``compute_max_speed(u)`` returns cmax and ``evolve(u, u_eq, ...)`` advances the
state.  Arrays are float64 in the reference's Fortran layout u(1:4,1:nx,1:ny)
(numpy shape (4, nx, ny), order="F").

The HIP library is the only compute path: if libfvwb.so is missing, loading
fails loudly; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libfvwb.so")
HEADER = os.path.join(os.path.dirname(PKG_DIR), "include", "fvwb.h")

OK, ERR_ARG, ERR_HIP, ERR_OOM, ERR_NONPHYSICAL, ERR_STATE = 0, -1, -2, -3, -4, -5
BC_FROZEN_RING, BC_REFLECT = 0, 1
RECON_FIRST_ORDER, RECON_MUSCL_MINMOD = 0, 1
FLUX_RUSANOV, FLUX_HLL = 0, 1
RK2_HEUN, SSPRK3 = 0, 1
FP64, FP32 = 0, 1

GAMMA_REF = float(np.float32(1.4))   # parameters_2d.f90:19 (default-real literal)
RHO0_REF = float(np.float32(1.21))   # benchmark_2d.f90:196


class FvwbError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fvwb error {code}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    """``fvwb_params`` of include/fvwb.h."""

    _fields_ = [
        ("nvar", ctypes.c_int32), ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("bc", ctypes.c_int32), ("recon", ctypes.c_int32), ("flux", ctypes.c_int32),
        ("integrator", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("gamma", ctypes.c_double), ("boxlen_x", ctypes.c_double),
        ("boxlen_y", ctypes.c_double), ("cfl", ctypes.c_double),
        ("rho0", ctypes.c_double), ("p0", ctypes.c_double), ("g", ctypes.c_double),
    ]


def make_params(nx, ny, *, bc=BC_FROZEN_RING, recon=RECON_FIRST_ORDER, flux=FLUX_RUSANOV,
                integrator=RK2_HEUN, precision=FP64, gamma=GAMMA_REF, boxlen_x=1.0, boxlen_y=1.0,
                cfl=0.5, rho0=RHO0_REF, p0=1.0, g=1.0) -> Params:
    return Params(4, nx, ny, bc, recon, flux, integrator, precision, gamma, boxlen_x, boxlen_y,
                  cfl, rho0, p0, g)


_lib = None


def declared_symbols(header=HEADER):
    """Names of the functions include/fvwb.h declares."""
    import re
    text = open(header).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(fvwb_\w+)\s*\(", text, re.M)))


def lib():
    """Load libfvwb.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FvwbError(ERR_STATE, f"{LIB_PATH} not built; run make -C fvm-source-wb_amd/csrc")
    L = ctypes.CDLL(LIB_PATH)
    H = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int32)
    pp = ctypes.POINTER(Params)
    sig = {
        "fvwb_params_reference_2d": [pp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32],
        "fvwb_create": [ctypes.POINTER(H), pp],
        "fvwb_destroy": [H],
        "fvwb_compute_update_exact": [H, dp, dp, dp],
        "fvwb_compute_max_speed": [H, dp, dp],
        "fvwb_upload": [H, dp, dp],
        "fvwb_init_gaussian": [H, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double],
        "fvwb_advance": [H, ctypes.c_int32, ctypes.c_double, dp, ip],
        "fvwb_advance_async": [H, ctypes.c_int32, ctypes.c_double],
        "fvwb_download": [H, dp],
        "fvwb_download_deviation": [H, dp],
        "fvwb_get_time": [H, dp, dp, ip],
        "fvwb_get_stream": [H, ctypes.POINTER(ctypes.c_void_p)],
        "fvwb_copy_bandwidth": [ctypes.c_uint64, ctypes.c_int32, dp],
        "fvwb_set_launch": [H, ctypes.c_int32, ctypes.c_int32],
        "fvwb_batch_create": [ctypes.POINTER(H), ctypes.c_int32, pp],
        "fvwb_batch_destroy": [H],
        "fvwb_batch_init_gaussian": [H, dp, ctypes.c_double, ctypes.c_double, ctypes.c_double],
        "fvwb_batch_upload": [H, ctypes.c_int32, dp, dp],
        "fvwb_batch_reset_time": [H],
        "fvwb_batch_advance": [H, ctypes.c_int32, ctypes.c_double],
        "fvwb_batch_advance_async": [H, ctypes.c_int32, ctypes.c_double],
        "fvwb_batch_download": [H, ctypes.c_int32, dp],
        "fvwb_batch_download_deviation": [H, ctypes.c_int32, dp],
        "fvwb_batch_get_time": [H, ctypes.c_int32, dp, dp, ip],
        "fvwb_batch_get_stream": [H, ctypes.POINTER(ctypes.c_void_p)],
        "fvwb_batch_use_graph": [H, ctypes.c_int32],
        "fvwb_batch_set_launch": [H, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32],
        "fvwb_batch_get_launch": [H, ip, ip, ip, ip],
        "fvwb_batch_autotune": [H, ctypes.c_int32, ctypes.c_char_p, ip, ip],
        "fvwb_get_launch": [H, ip, ip],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.fvwb_last_error.argtypes = [H]
    L.fvwb_last_error.restype = ctypes.c_char_p
    L.fvwb_batch_last_error.argtypes = [H]
    L.fvwb_batch_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def _check(rc, h=None):
    if rc != OK:
        msg = lib().fvwb_last_error(h)
        raise FvwbError(rc, msg.decode() if msg else "")


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _fortran(a, shape):
    a = np.asfortranarray(a, dtype=np.float64)
    if a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def reference_params(nx=300, ny=300, nequilibrium=2) -> Params:
    """The benchmark_2d set-up of parameters_2d.f90 (first-order Rusanov, frozen
    ring, Heun, FP64) at nx x ny cells."""
    p = Params()
    _check(lib().fvwb_params_reference_2d(ctypes.byref(p), nx, ny, nequilibrium))
    return p


class Solver2D:
    """One 2D instance on the current GPU (a C-ABI handle)."""

    def __init__(self, params: Params):
        self.params = params
        self.shape = (4, params.nx, params.ny)
        h = ctypes.c_void_p()
        _check(lib().fvwb_create(ctypes.byref(h), ctypes.byref(params)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().fvwb_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- reference external procedures ------------------------------------------
    def compute_update_exact(self, u, w_eq):
        """dudt of compute_update_exact(u, w_eq, dudt) (benchmark_2d.f90:465-618)."""
        u, w_eq = _fortran(u, self.shape), _fortran(w_eq, self.shape)
        out = np.zeros(self.shape, order="F")
        _check(lib().fvwb_compute_update_exact(self._h, _dp(u), _dp(w_eq), _dp(out)), self._h)
        return out

    def compute_max_speed(self, u):
        """cmax of compute_max_speed(u, cmax) (benchmark_2d.f90:264-279)."""
        u = _fortran(u, self.shape)
        c = ctypes.c_double()
        _check(lib().fvwb_compute_max_speed(self._h, _dp(u), ctypes.byref(c)), self._h)
        return c.value

    # ---- device-resident stepping -----------------------------------------------
    def upload(self, u, w_eq):
        u, w_eq = _fortran(u, self.shape), _fortran(w_eq, self.shape)
        _check(lib().fvwb_upload(self._h, _dp(u), _dp(w_eq)), self._h)

    def init_gaussian(self, amp=1e-3, x0=0.3, y0=0.3, width=100.0):
        _check(lib().fvwb_init_gaussian(self._h, amp, x0, y0, width), self._h)

    def advance(self, nsteps, t_end=math.inf):
        t = ctypes.c_double()
        n = ctypes.c_int32()
        _check(lib().fvwb_advance(self._h, nsteps, t_end, ctypes.byref(t), ctypes.byref(n)), self._h)
        return t.value, n.value

    def advance_async(self, nsteps, t_end=math.inf):
        _check(lib().fvwb_advance_async(self._h, nsteps, t_end), self._h)

    def time(self):
        t, dt = ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_int32()
        _check(lib().fvwb_get_time(self._h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(n)), self._h)
        return t.value, dt.value, n.value

    def download(self):
        out = np.zeros(self.shape, order="F")
        _check(lib().fvwb_download(self._h, _dp(out)), self._h)
        return out

    def download_deviation(self):
        out = np.zeros(self.shape, order="F")
        _check(lib().fvwb_download_deviation(self._h, _dp(out)), self._h)
        return out

    def set_launch(self, rows_per_strip, columns_per_lane=4):
        _check(lib().fvwb_set_launch(self._h, rows_per_strip, columns_per_lane), self._h)

    def get_launch(self):
        r, c = ctypes.c_int32(), ctypes.c_int32()
        _check(lib().fvwb_get_launch(self._h, ctypes.byref(r), ctypes.byref(c)), self._h)
        return r.value, c.value

    def stream(self):
        s = ctypes.c_void_p()
        _check(lib().fvwb_get_stream(self._h, ctypes.byref(s)), self._h)
        return s.value


def _bcheck(rc, h=None):
    if rc != OK:
        msg = lib().fvwb_batch_last_error(h)
        raise FvwbError(rc, msg.decode() if msg else "")


class Batch2D:
    """Many independent 2D instances advanced together (one persistent launch
    per stage); each instance has its own size, gravity, dt and clock."""

    def __init__(self, params_list):
        self.params = list(params_list)
        self.n = len(self.params)
        arr = (Params * self.n)(*self.params)
        h = ctypes.c_void_p()
        _bcheck(lib().fvwb_batch_create(ctypes.byref(h), self.n, arr))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().fvwb_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def shape(self, b):
        if not 0 <= b < self.n:
            raise FvwbError(ERR_ARG, f"instance index {b} out of range [0, {self.n})")
        return (4, self.params[b].nx, self.params[b].ny)

    def init_gaussian(self, amps, x0=0.3, y0=0.3, width=100.0):
        a = np.ascontiguousarray(amps, dtype=np.float64)
        assert a.shape == (self.n,)
        _bcheck(lib().fvwb_batch_init_gaussian(self._h, _dp(a), x0, y0, width), self._h)

    def upload(self, b, u, w_eq):
        u, w_eq = _fortran(u, self.shape(b)), _fortran(w_eq, self.shape(b))
        _bcheck(lib().fvwb_batch_upload(self._h, b, _dp(u), _dp(w_eq)), self._h)

    def reset_time(self):
        _bcheck(lib().fvwb_batch_reset_time(self._h), self._h)

    def advance(self, nsteps, t_end=math.inf):
        _bcheck(lib().fvwb_batch_advance(self._h, nsteps, t_end), self._h)

    def advance_async(self, nsteps, t_end=math.inf):
        _bcheck(lib().fvwb_batch_advance_async(self._h, nsteps, t_end), self._h)

    def download(self, b):
        out = np.zeros(self.shape(b), order="F")
        _bcheck(lib().fvwb_batch_download(self._h, b, _dp(out)), self._h)
        return out

    def download_deviation(self, b):
        out = np.zeros(self.shape(b), order="F")
        _bcheck(lib().fvwb_batch_download_deviation(self._h, b, _dp(out)), self._h)
        return out

    def time(self, b):
        t, dt = ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_int32()
        _bcheck(lib().fvwb_batch_get_time(self._h, b, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(n)), self._h)
        return t.value, dt.value, n.value

    def stream(self):
        s = ctypes.c_void_p()
        _bcheck(lib().fvwb_batch_get_stream(self._h, ctypes.byref(s)), self._h)
        return s.value

    def use_graph(self, enable=True):
        _bcheck(lib().fvwb_batch_use_graph(self._h, 1 if enable else 0), self._h)

    def set_launch(self, rows, cpl=2, blocks=0):
        _bcheck(lib().fvwb_batch_set_launch(self._h, rows, cpl, blocks), self._h)

    def get_launch(self):
        r, c, b, t = (ctypes.c_int32() for _ in range(4))
        _bcheck(lib().fvwb_batch_get_launch(self._h, ctypes.byref(r), ctypes.byref(c), ctypes.byref(b),
                                            ctypes.byref(t)), self._h)
        return r.value, c.value, b.value, t.value

    def autotune(self, trial_steps=3, cache_path=None):
        r, c = ctypes.c_int32(), ctypes.c_int32()
        path = cache_path.encode() if cache_path else None
        _bcheck(lib().fvwb_batch_autotune(self._h, trial_steps, path, ctypes.byref(r), ctypes.byref(c)), self._h)
        return r.value, c.value


def copy_bandwidth(nbytes=1 << 30, reps=10):
    """Device-to-device streaming copy bandwidth in GB/s (read + write)."""
    g = ctypes.c_double()
    _check(lib().fvwb_copy_bandwidth(nbytes, reps, ctypes.byref(g)))
    return g.value


def evolve(u, u_eq, params: Params, t_end=math.inf, max_steps=1 << 30):
    """evolve(u, u_eq) of benchmark_2d.f90:221-260 on the GPU: returns the new
    state, the final time and the number of steps (u_eq is the primitive
    equilibrium, as at the reference's call site benchmark_2d.f90:16-18)."""
    with Solver2D(params) as s:
        s.upload(u, u_eq)
        t, n = s.advance(max_steps, t_end)
        return s.download(), t, n
