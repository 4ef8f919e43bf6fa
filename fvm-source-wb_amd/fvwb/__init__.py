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
FLUX_RUSANOV, FLUX_HLL, FLUX_HLLC = 0, 1, 2
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


BC1D_PERIODIC, BC1D_ZERO_GRADIENT, BC1D_REFLEXIVE, BC1D_WALL = 1, 2, 3, 4


class Params1D(ctypes.Structure):
    """``fvwb1d_params`` of include/fvwb.h."""

    _fields_ = [
        ("nvar", ctypes.c_int32), ("nx", ctypes.c_int32), ("bc", ctypes.c_int32),
        ("recon", ctypes.c_int32), ("integrator", ctypes.c_int32), ("nequilibrium", ctypes.c_int32),
        ("gamma", ctypes.c_double), ("boxlen", ctypes.c_double), ("cfl", ctypes.c_double),
        ("g", ctypes.c_double),
    ]


def make_params_1d(nx, *, bc=BC1D_WALL, recon=RECON_MUSCL_MINMOD, integrator=RK2_HEUN, nequilibrium=2,
                   gamma=1.4, boxlen=1.0, cfl=0.5, g=1.0) -> Params1D:
    """Defaults: the 1D gravity workload (MUSCL-minmod, Rusanov, SSP-RK2, walls, CFL 0.5)."""
    return Params1D(3, nx, bc, recon, integrator, nequilibrium, gamma, boxlen, cfl, g)


DG_LIMITER_NONE, DG_LIMITER_TVB = 0, 1
DG_SSPRK3, DG_SSPRK54 = 0, 1


class DGParams(ctypes.Structure):
    """``fvwb_dg_params`` of include/fvwb.h."""

    _fields_ = [
        ("nvar", ctypes.c_int32), ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("degree", ctypes.c_int32), ("flux", ctypes.c_int32), ("limiter", ctypes.c_int32),
        ("bc", ctypes.c_int32), ("integrator", ctypes.c_int32),
        ("gamma", ctypes.c_double), ("boxlen_x", ctypes.c_double), ("boxlen_y", ctypes.c_double),
        ("cfl", ctypes.c_double), ("rho0", ctypes.c_double), ("p0", ctypes.c_double),
        ("g", ctypes.c_double), ("tvb_m", ctypes.c_double),
    ]


DG1D_RIEMANN_LLF, DG1D_RIEMANN_HLLC = 1, 2


class DG1DParams(ctypes.Structure):
    """``fvwb_dg1d_params`` of include/fvwb.h."""

    _fields_ = [
        ("nvar", ctypes.c_int32), ("nx", ctypes.c_int32), ("n", ctypes.c_int32),
        ("riemann", ctypes.c_int32), ("source", ctypes.c_int32), ("pad", ctypes.c_int32),
        ("gamma", ctypes.c_double), ("boxlen", ctypes.c_double),
    ]


class RDParams(ctypes.Structure):
    """``fvwb_rd_params`` of include/fvwb.h (benchmark_rd.f90's parameters_rd)."""

    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("bc", ctypes.c_int32), ("pad", ctypes.c_int32),
        ("gamma", ctypes.c_double), ("boxlen_x", ctypes.c_double), ("boxlen_y", ctypes.c_double),
        ("cfl", ctypes.c_double),
    ]


RD_BC_PERIODIC, RD_BC_TRANSMISSIVE = 2, 3


class DGQParams(ctypes.Structure):
    """``fvwb_dgq_params`` of include/fvwb.h: the compile-time switches of
    2d/parameters_dg_2d.f90 for the reference's own 2D DG program."""

    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("order", ctypes.c_int32),
        ("flux_type", ctypes.c_int32), ("source", ctypes.c_int32), ("grad_phi_case", ctypes.c_int32),
        ("bc", ctypes.c_int32), ("ninit", ctypes.c_int32), ("solver", ctypes.c_int32),
        ("use_limiter", ctypes.c_int32), ("gamma", ctypes.c_double), ("boxlen_x", ctypes.c_double),
        ("boxlen_y", ctypes.c_double), ("cfl", ctypes.c_double), ("eps", ctypes.c_double),
    ]


DGQ_FLUX_NONE, DGQ_FLUX_LLF1, DGQ_FLUX_HLL2, DGQ_FLUX_HLLC = 0, 1, 2, 3
DGQ_SOLVER_RK4, DGQ_SOLVER_SS4, DGQ_SOLVER_EQL, DGQ_SOLVER_DEB = 0, 1, 2, 3


def reference_params_dgq(n=8, **changes) -> DGQParams:
    """The shipped 2d/parameters_dg_2d.f90 set-up (8x8, mx = my = 2, 'llf',
    source 1, grad_phi_case 2, bc 1, ninit 1, 'RK4', 'ONP'; gamma, cfl and eps
    as default-real literals) at n x n cells, with `changes` applied."""
    p = DGQParams()
    _checkdgq(lib().fvwb_dgq_params_reference(ctypes.byref(p)))
    p.nx = p.ny = n
    for k, v in changes.items():
        setattr(p, k, v)
    return p


def make_rd_params(nx, ny, *, bc=RD_BC_TRANSMISSIVE, gamma=float(np.float32(1.4)), boxlen_x=1.0, boxlen_y=1.0,
                   cfl=0.5) -> RDParams:
    """gamma defaults to the reference's default-real 1.4, as its other parameter modules set it."""
    return RDParams(nx, ny, bc, 0, gamma, boxlen_x, boxlen_y, cfl)


def make_dg_params(nx, ny, *, degree=2, flux=FLUX_HLL, limiter=DG_LIMITER_TVB, gamma=1.4, boxlen_x=1.0,
                   boxlen_y=1.0, cfl=0.8, rho0=1.0, p0=1.0, g=1.0, tvb_m=50.0, integrator=0) -> DGParams:
    """Defaults: the DG gravity workload (P2, HLL, TVB with M = 50, walls, CFL 0.8)."""
    return DGParams(4, nx, ny, degree, flux, limiter, BC_REFLECT, integrator, gamma, boxlen_x, boxlen_y, cfl,
                    rho0, p0, g, tvb_m)


_lib = None


def declared_symbols(header=HEADER):
    """Names of the functions include/fvwb.h declares."""
    import re
    text = open(header).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(fvwb(?:1d)?_\w+)\s*\(", text, re.M)))


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
    cs = ctypes.c_char_p
    sig = {
        "fvwb_checkpoint_save": [H, cs], "fvwb_checkpoint_load": [H, cs],
        "fvwb1d_checkpoint_save": [H, cs], "fvwb1d_checkpoint_load": [H, cs],
        "fvwb_dg_checkpoint_save": [H, cs], "fvwb_dg_checkpoint_load": [H, cs],
        "fvwb_dg1d_checkpoint_save": [H, cs], "fvwb_dg1d_checkpoint_load": [H, cs],
        "fvwb_rd_checkpoint_save": [H, cs], "fvwb_rd_checkpoint_load": [H, cs],
        "fvwb_params_reference_2d": [pp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32],
        "fvwb_create": [ctypes.POINTER(H), pp],
        "fvwb_destroy": [H],
        "fvwb_compute_update_exact": [H, dp, dp, dp],
        "fvwb_compute_update": [H, dp, dp, dp],
        "fvwb_compute_residual": [H, dp, dp, dp],
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
    pdg = ctypes.POINTER(DGParams)
    sig.update({
        "fvwb_dg_create": [ctypes.POINTER(H), pdg],
        "fvwb_dg_destroy": [H],
        "fvwb_dg_compute_residual": [H, dp, dp],
        "fvwb_dg_upload": [H, dp],
        "fvwb_dg_init_fourier": [H, ctypes.c_int32, dp, dp, dp, dp, ctypes.c_double],
        "fvwb_dg_advance": [H, ctypes.c_int32, ctypes.c_double],
        "fvwb_dg_advance_async": [H, ctypes.c_int32, ctypes.c_double],
        "fvwb_dg_download": [H, dp],
        "fvwb_dg_get_time": [H, dp, dp, ip],
        "fvwb_dg_get_stream": [H, ctypes.POINTER(ctypes.c_void_p)],
    })
    pd1 = ctypes.POINTER(DG1DParams)
    sig.update({
        "fvwb_dg1d_params_reference": [pd1, ctypes.c_int32],
        "fvwb_dg1d_create": [ctypes.POINTER(H), pd1],
        "fvwb_dg1d_destroy": [H],
        "fvwb_dg1d_compute_update_exact_delta": [H, dp, dp, dp],
        "fvwb_dg1d_upload": [H, dp, dp, dp],
        "fvwb_dg1d_advance": [H, ctypes.c_int32, ctypes.c_double],
        "fvwb_dg1d_download": [H, dp, dp],
        "fvwb_dg1d_get_time": [H, dp, dp, ip],
    })
    prd = ctypes.POINTER(RDParams)
    sig.update({
        "fvwb_rd_create": [ctypes.POINTER(H), prd],
        "fvwb_rd_destroy": [H],
        "fvwb_rd_compute_update": [H, dp, dp, ctypes.c_double, dp],
        "fvwb_rd_compute_max_speed": [H, dp, dp],
        "fvwb_rd_upload": [H, dp],
        "fvwb_rd_advance": [H, ctypes.c_int32, ctypes.c_double],
        "fvwb_rd_download": [H, dp],
        "fvwb_rd_get_time": [H, dp, dp, ctypes.POINTER(ctypes.c_int64)],
    })
    pq = ctypes.POINTER(DGQParams)
    sig.update({
        "fvwb_dgq_params_reference": [pq],
        "fvwb_dgq_gl_quadrature": [ctypes.c_int32, dp, dp],
        "fvwb_dgq_create": [ctypes.POINTER(H), pq],
        "fvwb_dgq_destroy": [H],
        "fvwb_dgq_compute_update": [H, dp, dp, dp, dp, dp],
        "fvwb_dgq_compute_max_speed": [H, dp, dp, dp, dp, dp],
        "fvwb_dgq_apply_limiter": [H, dp],
        "fvwb_dgq_modes_from_nodes": [H, dp, dp],
        "fvwb_dgq_nodes_from_modes": [H, dp, dp],
        "fvwb_dgq_upload": [H, dp, dp, dp],
        "fvwb_dgq_advance": [H, ctypes.c_int32, ctypes.c_double],
        "fvwb_dgq_download": [H, dp],
        "fvwb_dgq_get_time": [H, dp, dp, ctypes.POINTER(ctypes.c_int64)],
        "fvwb_dgq_compute_error": [H, dp, dp, dp, dp, dp],
        "fvwb_dgq_compute_error_noalloc": [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    })
    p1 = ctypes.POINTER(Params1D)
    sig.update({
        "fvwb1d_params_reference": [p1, ctypes.c_int32, ctypes.c_int32],
        "fvwb1d_create": [ctypes.POINTER(H), ctypes.c_int32, p1],
        "fvwb1d_destroy": [H],
        "fvwb1d_compute_update": [H, dp, dp, dp],
        "fvwb1d_compute_update_sr": [H, dp, dp],
        "fvwb1d_compute_update_fvm": [H, dp, dp],
        "fvwb1d_fvm_compute_update": [H, ctypes.c_int32, dp, dp],
        "fvwb1d_compute_max_speed": [H, dp, dp],
        "fvwb1d_compute_residual": [H, ctypes.c_int32, dp, dp, dp],
        "fvwb1d_upload": [H, ctypes.c_int32, dp, dp],
        "fvwb1d_init_gaussian": [H, dp, ctypes.c_double, ctypes.c_double],
        "fvwb1d_advance": [H, ctypes.c_int32, ctypes.c_double],
        "fvwb1d_download": [H, ctypes.c_int32, dp],
        "fvwb1d_get_time": [H, ctypes.c_int32, dp, dp, ip],
        "fvwb1d_get_stream": [H, ctypes.POINTER(ctypes.c_void_p)],
    })
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.fvwb_last_error.argtypes = [H]
    L.fvwb_last_error.restype = ctypes.c_char_p
    L.fvwb_batch_last_error.argtypes = [H]
    L.fvwb_batch_last_error.restype = ctypes.c_char_p
    L.fvwb1d_last_error.argtypes = [H]
    L.fvwb1d_last_error.restype = ctypes.c_char_p
    L.fvwb_dg_last_error.argtypes = [H]
    L.fvwb_dg_last_error.restype = ctypes.c_char_p
    L.fvwb_dg1d_last_error.argtypes = [H]
    L.fvwb_dg1d_last_error.restype = ctypes.c_char_p
    L.fvwb_rd_last_error.argtypes = [H]
    L.fvwb_rd_last_error.restype = ctypes.c_char_p
    L.fvwb_dgq_last_error.argtypes = [H]
    L.fvwb_dgq_last_error.restype = ctypes.c_char_p
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
    def checkpoint_save(self, path):
        """Write the device state (deviation at t^n, equilibrium, clock) to path."""
        _check(lib().fvwb_checkpoint_save(self._h, os.fsencode(path)), self._h)

    def checkpoint_load(self, path):
        """Restore a state written by checkpoint_save of a handle with the same parameters."""
        _check(lib().fvwb_checkpoint_load(self._h, os.fsencode(path)), self._h)

    def compute_update_exact(self, u, w_eq):
        """dudt of compute_update_exact(u, w_eq, dudt) (benchmark_2d.f90:465-618)."""
        u, w_eq = _fortran(u, self.shape), _fortran(w_eq, self.shape)
        out = np.zeros(self.shape, order="F")
        _check(lib().fvwb_compute_update_exact(self._h, _dp(u), _dp(w_eq), _dp(out)), self._h)
        return out

    def compute_update(self, u, w_eq=None):
        """dudt of compute_update(u, w_eq, dudt) (benchmark_2d.f90:370-463), the plain non-WB residual."""
        u = _fortran(u, self.shape)
        w_eq = np.zeros(self.shape, order="F") if w_eq is None else _fortran(w_eq, self.shape)
        out = np.zeros(self.shape, order="F")
        _check(lib().fvwb_compute_update(self._h, _dp(u), _dp(w_eq), _dp(out)), self._h)
        return out

    def compute_residual(self, u, w_eq):
        """L(u) of the configured scheme, the residual advance() integrates (y faces at (j-1)*dy)."""
        u, w_eq = _fortran(u, self.shape), _fortran(w_eq, self.shape)
        out = np.zeros(self.shape, order="F")
        _check(lib().fvwb_compute_residual(self._h, _dp(u), _dp(w_eq), _dp(out)), self._h)
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


# ---- 1D (benchmark_1d.f90) ---------------------------------------------------------
def reference_params_1d(nx, nequilibrium=2) -> Params1D:
    """benchmark_1d.f90's set-up (parameters.f90) for nx cells."""
    p = Params1D()
    rc = lib().fvwb1d_params_reference(ctypes.byref(p), nx, nequilibrium)
    if rc != OK:
        raise FvwbError(rc, "bad reference parameters")
    return p


def _check1d(rc, h=None):
    if rc != OK:
        msg = lib().fvwb1d_last_error(h)
        raise FvwbError(rc, msg.decode() if msg else "")


class Solver1D:
    """``ninst`` independent 1D problems on the GPU (instance 0 serves the
    reference's per-stage procedures).  Arrays are float64 AoS(3,nx):
    numpy shape (3, nx), order="F"."""

    def __init__(self, params):
        if isinstance(params, Params1D):
            params = [params]
        self.params = list(params)
        arr = (Params1D * len(self.params))(*self.params)
        self._h = ctypes.c_void_p()
        _check1d(lib().fvwb1d_create(ctypes.byref(self._h), len(self.params), arr))

    def close(self):
        if self._h:
            lib().fvwb1d_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _nx(self, inst):
        if not 0 <= inst < len(self.params):
            raise FvwbError(ERR_ARG, f"instance {inst} out of range")
        return self.params[inst].nx

    def _arr(self, a, inst=0):
        a = np.asfortranarray(a, dtype=np.float64)
        if a.shape != (3, self._nx(inst)):
            raise FvwbError(ERR_ARG, f"expected shape (3, {self._nx(inst)}), got {a.shape}")
        return a

    def checkpoint_save(self, path):
        _check1d(lib().fvwb1d_checkpoint_save(self._h, os.fsencode(path)), self._h)

    def checkpoint_load(self, path):
        _check1d(lib().fvwb1d_checkpoint_load(self._h, os.fsencode(path)), self._h)

    def compute_update(self, u, w_eq):
        """dudt of compute_update(u, w_eq, dudt) (benchmark_1d.f90:263-377), instance 0."""
        u, w_eq = self._arr(u), self._arr(w_eq)
        out = np.zeros((3, self._nx(0)), order="F")
        _check1d(lib().fvwb1d_compute_update(self._h, _dp(u), _dp(w_eq), _dp(out)), self._h)
        return out

    def compute_update_sr(self, u, w_eq=None):
        """dudt of compute_update_sr(u, w_eq, dudt) (benchmark_1d.f90:553-747), the
        hydrostatic-reconstruction solver 'WB1', instance 0; w_eq is unused there."""
        u = self._arr(u)
        out = np.zeros((3, self._nx(0)), order="F")
        _check1d(lib().fvwb1d_compute_update_sr(self._h, _dp(u), _dp(out)), self._h)
        return out

    def compute_update_fvm(self, u, w_eq=None):
        """dudt of compute_update_fvm(u, w_eq, dudt) (benchmark_1d.f90:454-549), the
        plain solver 'FVM', instance 0; w_eq is unused there."""
        u = self._arr(u)
        out = np.zeros((3, self._nx(0)), order="F")
        _check1d(lib().fvwb1d_compute_update_fvm(self._h, _dp(u), _dp(out)), self._h)
        return out

    def fvm_compute_update(self, u, source=2):
        """dudt of compute_update(u, dudt) of fvm.f90 (:188-251), instance 0, with
        the handle's bc (1 or 2) and fvm_commons' source switch (1 none, 2 gravity)."""
        u = self._arr(u)
        out = np.zeros((3, self._nx(0)), order="F")
        _check1d(lib().fvwb1d_fvm_compute_update(self._h, int(source), _dp(u), _dp(out)), self._h)
        return out

    def compute_max_speed(self, u):
        """cmax of compute_max_speed(u, cmax) (benchmark_1d.f90:157-169), instance 0."""
        u = self._arr(u)
        c = ctypes.c_double()
        _check1d(lib().fvwb1d_compute_max_speed(self._h, _dp(u), ctypes.byref(c)), self._h)
        return c.value

    def compute_residual(self, u, w_eq, inst=0):
        u, w_eq = self._arr(u, inst), self._arr(w_eq, inst)
        out = np.zeros((3, self._nx(inst)), order="F")
        _check1d(lib().fvwb1d_compute_residual(self._h, inst, _dp(u), _dp(w_eq), _dp(out)), self._h)
        return out

    def upload(self, u, w_eq, inst=0):
        u, w_eq = self._arr(u, inst), self._arr(w_eq, inst)
        _check1d(lib().fvwb1d_upload(self._h, inst, _dp(u), _dp(w_eq)), self._h)

    def init_gaussian(self, amp, x0=0.5, width=0.05):
        amp = np.ascontiguousarray(np.broadcast_to(np.asarray(amp, dtype=np.float64), (len(self.params),)))
        _check1d(lib().fvwb1d_init_gaussian(self._h, _dp(amp), x0, width), self._h)

    def advance(self, nsteps=None, t_end=math.inf):
        """Advance every instance until t_end (last step clipped) or nsteps more steps."""
        if nsteps is None:
            if not math.isfinite(t_end):
                raise FvwbError(ERR_ARG, "give nsteps or a finite t_end")
            nsteps = 2**31 - 1
        _check1d(lib().fvwb1d_advance(self._h, int(min(nsteps, 2**31 - 1)), float(t_end)), self._h)

    def download(self, inst=0):
        out = np.zeros((3, self._nx(inst)), order="F")
        _check1d(lib().fvwb1d_download(self._h, inst, _dp(out)), self._h)
        return out

    def time(self, inst=0):
        t, dt = ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_int32()
        self._nx(inst)
        _check1d(lib().fvwb1d_get_time(self._h, inst, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(n)),
                 self._h)
        return t.value, dt.value, n.value

    def stream(self):
        s = ctypes.c_void_p()
        _check1d(lib().fvwb1d_get_stream(self._h, ctypes.byref(s)), self._h)
        return s.value or 0


# ---- 2D modal DG ---------------------------------------------------------------------
def dg_fourier_terms(n=16, seed=0xF00D, mmax=8):
    """(a_k, m_k, n_k, theta_k) of the DG workload's pressure perturbation:
    a ~ U[0.5,1], (m,n) in {1..mmax}^2, theta ~ U[0, 2 pi), drawn from numpy's
    PCG64 seeded with `seed` (the problem definition, not a solver input)."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.5, 1.0, n)
    mk = rng.integers(1, mmax + 1, n)
    nk = rng.integers(1, mmax + 1, n)
    th = rng.uniform(0.0, 2 * np.pi, n)
    return a, mk, nk, th


def _checkdg(rc, h=None):
    if rc != OK:
        msg = lib().fvwb_dg_last_error(h)
        raise FvwbError(rc, msg.decode() if msg else "")


class SolverDG:
    """P2 modal DG of the deviation on the GPU (FP32, projections on MFMA).
    Modes are float64 AoS(4,nx,ny,6): numpy shape (4, nx, ny, 6), order="F"."""

    NM = 6

    def __init__(self, params: DGParams):
        self.params = params
        self.shape = (4, params.nx, params.ny, self.NM)
        self._h = ctypes.c_void_p()
        _checkdg(lib().fvwb_dg_create(ctypes.byref(self._h), ctypes.byref(params)))

    def close(self):
        if self._h:
            lib().fvwb_dg_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _arr(self, a):
        a = np.asfortranarray(a, dtype=np.float64)
        if a.shape != self.shape:
            raise FvwbError(ERR_ARG, f"expected shape {self.shape}, got {a.shape}")
        return a

    def compute_residual(self, modes):
        m = self._arr(modes)
        out = np.zeros(self.shape, order="F")
        _checkdg(lib().fvwb_dg_compute_residual(self._h, _dp(m), _dp(out)), self._h)
        return out

    def checkpoint_save(self, path):
        _checkdg(lib().fvwb_dg_checkpoint_save(self._h, os.fsencode(path)), self._h)

    def checkpoint_load(self, path):
        _checkdg(lib().fvwb_dg_checkpoint_load(self._h, os.fsencode(path)), self._h)

    def upload(self, modes):
        m = self._arr(modes)
        _checkdg(lib().fvwb_dg_upload(self._h, _dp(m)), self._h)

    def init_fourier(self, a, mk, nk, theta, amp):
        arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (a, mk, nk, theta)]
        _checkdg(lib().fvwb_dg_init_fourier(self._h, len(arrs[0]), *(_dp(x) for x in arrs), float(amp)), self._h)

    def advance(self, nsteps, t_end=math.inf):
        _checkdg(lib().fvwb_dg_advance(self._h, int(nsteps), float(t_end)), self._h)

    def advance_async(self, nsteps, t_end=math.inf):
        _checkdg(lib().fvwb_dg_advance_async(self._h, int(nsteps), float(t_end)), self._h)

    def download(self):
        out = np.zeros(self.shape, order="F")
        _checkdg(lib().fvwb_dg_download(self._h, _dp(out)), self._h)
        return out

    def time(self):
        t, dt = ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_int32()
        _checkdg(lib().fvwb_dg_get_time(self._h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(n)), self._h)
        return t.value, dt.value, n.value

    def stream(self):
        s = ctypes.c_void_p()
        _checkdg(lib().fvwb_dg_get_stream(self._h, ctypes.byref(s)), self._h)
        return s.value or 0


# ---- 1D modal DG (dg_with_source.f90) ------------------------------------------------
def reference_params_dg1d(nx=128) -> DG1DParams:
    """dg_commons.f90's set-up: n = 3 modes, HLLC, gravity source, gamma 1.4 (FP32)."""
    p = DG1DParams()
    rc = lib().fvwb_dg1d_params_reference(ctypes.byref(p), nx)
    if rc != OK:
        raise FvwbError(rc, "bad reference parameters")
    return p


def _checkdg1d(rc, h=None):
    if rc != OK:
        msg = lib().fvwb_dg1d_last_error(h)
        raise FvwbError(rc, msg.decode() if msg else "")


class SolverDG1D:
    """The reference's 1D modal DG on the GPU (FP64).  Modes and the equilibrium
    at the Gauss points are float64 AoS(3,n,nx): numpy shape (3, n, nx), order="F"."""

    def __init__(self, params: DG1DParams):
        self.params = params
        self.shape = (3, params.n, params.nx)
        self._h = ctypes.c_void_p()
        _checkdg1d(lib().fvwb_dg1d_create(ctypes.byref(self._h), ctypes.byref(params)))

    def close(self):
        if self._h:
            lib().fvwb_dg1d_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _arr(self, a):
        a = np.asfortranarray(a, dtype=np.float64)
        if a.shape != self.shape:
            raise FvwbError(ERR_ARG, f"expected shape {self.shape}, got {a.shape}")
        return a

    def compute_update_exact_delta(self, delta_u, u_eq):
        """dudt of compute_update_exact_delta(delta_u, u_eq, dudt) (dg_with_source.f90:1749-2031)."""
        d, e = self._arr(delta_u), self._arr(u_eq)
        out = np.zeros(self.shape, order="F")
        _checkdg1d(lib().fvwb_dg1d_compute_update_exact_delta(self._h, _dp(d), _dp(e), _dp(out)), self._h)
        return out

    def upload(self, delta_u, u_eq, u_init):
        """Modes, equilibrium and the nodal state of the first dt, all AoS(3,n,nx)."""
        d, e, u = self._arr(delta_u), self._arr(u_eq), self._arr(u_init)
        _checkdg1d(lib().fvwb_dg1d_upload(self._h, _dp(d), _dp(e), _dp(u)), self._h)

    def advance(self, nsteps=(1 << 31) - 1, t_end=math.inf):
        """The main program's 'RKi' loop while t < t_end, at most nsteps steps."""
        _checkdg1d(lib().fvwb_dg1d_advance(self._h, int(nsteps), float(t_end)), self._h)

    def download(self):
        """(modes, u_init) after the last step."""
        d, u = np.zeros(self.shape, order="F"), np.zeros(self.shape, order="F")
        _checkdg1d(lib().fvwb_dg1d_download(self._h, _dp(d), _dp(u)), self._h)
        return d, u

    def time(self):
        t, dt = ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_int32()
        _checkdg1d(lib().fvwb_dg1d_get_time(self._h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(n)), self._h)
        return t.value, dt.value, n.value

    def checkpoint_save(self, path):
        _checkdg1d(lib().fvwb_dg1d_checkpoint_save(self._h, os.fsencode(path)), self._h)

    def checkpoint_load(self, path):
        _checkdg1d(lib().fvwb_dg1d_checkpoint_load(self._h, os.fsencode(path)), self._h)


def _checkrd(rc, h=None):
    if rc != OK:
        msg = lib().fvwb_rd_last_error(h)
        raise FvwbError(rc, msg.decode() if msg else "")


class SolverRD:
    """The reference's 2D residual distribution (benchmark_rd.f90, solver 'RD1')
    on the GPU (FP64).  States are float64 AoS(4,nx,ny): numpy shape (4, nx, ny),
    order="F", the reference's u(1:4,1:nx,1:ny) with mx = my = 1."""

    def __init__(self, params: RDParams):
        self.params = params
        self.shape = (4, params.nx, params.ny)
        self._h = ctypes.c_void_p()
        _checkrd(lib().fvwb_rd_create(ctypes.byref(self._h), ctypes.byref(params)))

    def close(self):
        if self._h:
            lib().fvwb_rd_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _arr(self, a):
        a = np.asfortranarray(a, dtype=np.float64)
        if a.shape != self.shape:
            raise FvwbError(ERR_ARG, f"expected shape {self.shape}, got {a.shape}")
        return a

    def compute_update(self, u, diff, dt):
        """dudt of compute_update_rd(u, diff, dt, x, y, dudt) (benchmark_rd.f90:612-745)."""
        a, d = self._arr(u), self._arr(diff)
        out = np.zeros(self.shape, order="F")
        _checkrd(lib().fvwb_rd_compute_update(self._h, _dp(a), _dp(d), float(dt), _dp(out)), self._h)
        return out

    def compute_max_speed(self, u):
        """compute_max_speed(u, speed_max) (benchmark_rd.f90:448-475)."""
        c = ctypes.c_double()
        _checkrd(lib().fvwb_rd_compute_max_speed(self._h, _dp(self._arr(u)), ctypes.byref(c)), self._h)
        return c.value

    def upload(self, u):
        _checkrd(lib().fvwb_rd_upload(self._h, _dp(self._arr(u))), self._h)

    def advance(self, nsteps=(1 << 31) - 1, t_end=math.inf):
        """evolve's loop (benchmark_rd.f90:356-385) while t < t_end, at most nsteps steps."""
        _checkrd(lib().fvwb_rd_advance(self._h, int(nsteps), float(t_end)), self._h)

    def download(self):
        out = np.zeros(self.shape, order="F")
        _checkrd(lib().fvwb_rd_download(self._h, _dp(out)), self._h)
        return out

    def checkpoint_save(self, path):
        """Write the resident state and clock to a binary checkpoint."""
        _checkrd(lib().fvwb_rd_checkpoint_save(self._h, os.fsencode(path)), self._h)

    def checkpoint_load(self, path):
        """Restore a state written by checkpoint_save of a handle with the same parameters."""
        _checkrd(lib().fvwb_rd_checkpoint_load(self._h, os.fsencode(path)), self._h)

    def time(self):
        t, dt = ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_int64()
        _checkrd(lib().fvwb_rd_get_time(self._h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(n)), self._h)
        return t.value, dt.value, n.value


def dgq_gl_quadrature(n):
    """gl_quadrature of 2d/legendre.f90:77-108 as the library computes it: (x, w)."""
    x, w = np.zeros(n), np.zeros(n)
    _checkdgq(lib().fvwb_dgq_gl_quadrature(int(n), _dp(x), _dp(w)))
    return x, w


def _checkdgq(rc, h=None):
    if rc != OK:
        msg = lib().fvwb_dgq_last_error(h)
        raise FvwbError(rc, msg.decode() if msg else "")


class SolverDGQ:
    """The reference's own 2D DG program (2d/benchmark_2d_dg.f90: tensor
    Legendre basis, mx = my = order, its flux dispatch, sources, positivity
    limiter and solvers) on the GPU in FP64, equal to the reference bit for bit.
    Modes and nodal values are float64 (4, nx, ny, order, order), order="F"
    (the reference's delta_u(1:4,1:nx,1:ny,1:mx,1:my)); coordinates
    (nx, ny, order, order)."""

    def __init__(self, params: DGQParams):
        self.params = params
        o = params.order
        self.shape = (4, params.nx, params.ny, o, o)
        self.cshape = (params.nx, params.ny, o, o)
        self._h = ctypes.c_void_p()
        _checkdgq(lib().fvwb_dgq_create(ctypes.byref(self._h), ctypes.byref(params)))

    def close(self):
        if self._h:
            lib().fvwb_dgq_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _arr(self, a, shape=None):
        shape = shape or self.shape
        a = np.asfortranarray(a, dtype=np.float64)
        if a.shape != shape:
            raise FvwbError(ERR_ARG, f"expected shape {shape}, got {a.shape}")
        return a

    def compute_update(self, delta_u, x, y, u_eq=None):
        """compute_update(delta_u, x, y, u_eq, dudt) (2d/benchmark_2d_dg.f90:1137-1479)."""
        a, xx, yy = self._arr(delta_u), self._arr(x, self.cshape), self._arr(y, self.cshape)
        out = np.zeros(self.shape, order="F")
        _checkdgq(lib().fvwb_dgq_compute_update(self._h, _dp(a), _dp(xx), _dp(yy), None, _dp(out)), self._h)
        return out

    def compute_max_speed(self, u):
        """compute_max_speed(u, cs_max, v_xmax, v_ymax, speed_max) (:826-870) of u(4,nx,ny)."""
        a = self._arr(u, (4, self.params.nx, self.params.ny))
        r = [ctypes.c_double() for _ in range(4)]
        _checkdgq(lib().fvwb_dgq_compute_max_speed(self._h, _dp(a), *[ctypes.byref(v) for v in r]), self._h)
        return tuple(v.value for v in r)

    def apply_limiter(self, u):
        """apply_limiter(u) (:1516-1555), 'ONP': returns the limited modes."""
        a = self._arr(u).copy(order="F")
        _checkdgq(lib().fvwb_dgq_apply_limiter(self._h, _dp(a)), self._h)
        return a

    def modes_from_nodes(self, nodes):
        out = np.zeros(self.shape, order="F")
        _checkdgq(lib().fvwb_dgq_modes_from_nodes(self._h, _dp(self._arr(nodes)), _dp(out)), self._h)
        return out

    def nodes_from_modes(self, modes):
        out = np.zeros(self.shape, order="F")
        _checkdgq(lib().fvwb_dgq_nodes_from_modes(self._h, _dp(self._arr(modes)), _dp(out)), self._h)
        return out

    def upload(self, delta_u, x, y):
        _checkdgq(lib().fvwb_dgq_upload(self._h, _dp(self._arr(delta_u)), _dp(self._arr(x, self.cshape)),
                                        _dp(self._arr(y, self.cshape))), self._h)

    def advance(self, nsteps=(1 << 31) - 1, t_end=math.inf):
        """evolve's loop (:666-769) while t < t_end, at most nsteps steps."""
        _checkdgq(lib().fvwb_dgq_advance(self._h, int(nsteps), float(t_end)), self._h)

    def download(self):
        out = np.zeros(self.shape, order="F")
        _checkdgq(lib().fvwb_dgq_download(self._h, _dp(out)), self._h)
        return out

    def time(self):
        t, dt = ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_int64()
        _checkdgq(lib().fvwb_dgq_get_time(self._h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(n)), self._h)
        return t.value, dt.value, n.value

    def compute_error(self, u_init, nodes=None):
        """compute_error (:23-89): (linf[4], l1[4], l2[4]) of nodes (default: the
        resident state's nodal values) against u_init."""
        ui = self._arr(u_init)
        nd = None if nodes is None else self._arr(nodes)
        linf, l1, l2 = np.zeros(4), np.zeros(4), np.zeros(4)
        _checkdgq(lib().fvwb_dgq_compute_error(self._h, None if nd is None else _dp(nd), _dp(ui), _dp(linf),
                                               _dp(l1), _dp(l2)), self._h)
        return linf, l1, l2
