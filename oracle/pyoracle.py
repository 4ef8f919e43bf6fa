"""ctypes binding of the CPU FP64 oracle (liboracle_fvwb.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline,
never by the product package.

Arrays are numpy float64 in the reference's Fortran layout AoS(4,nx,ny): use
``np.zeros((4, nx, ny), order="F")`` or equivalently a C array of shape
(ny, nx, 4).  The helpers below accept and return Fortran-ordered (4,nx,ny).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_fvwb.so")

BC_FROZEN_RING, BC_REFLECT = 0, 1
RECON_FIRST_ORDER, RECON_MUSCL_MINMOD = 0, 1
FLUX_RUSANOV, FLUX_HLL, FLUX_HLLC = 0, 1, 2
RK2_HEUN, SSPRK3 = 0, 1
FP64, FP32 = 0, 1

GAMMA_REF = float(np.float32(1.4))      # parameters_2d.f90:19, default-real literal
RHO0_REF = float(np.float32(1.21))      # benchmark_2d.f90:196


class Params(ctypes.Structure):
    """Mirror of ``fvwb_params`` (include/fvwb.h)."""

    _fields_ = [
        ("nvar", ctypes.c_int32), ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("bc", ctypes.c_int32), ("recon", ctypes.c_int32), ("flux", ctypes.c_int32),
        ("integrator", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("gamma", ctypes.c_double), ("boxlen_x", ctypes.c_double),
        ("boxlen_y", ctypes.c_double), ("cfl", ctypes.c_double),
        ("rho0", ctypes.c_double), ("p0", ctypes.c_double), ("g", ctypes.c_double),
    ]


BC1D_PERIODIC, BC1D_ZERO_GRADIENT, BC1D_REFLEXIVE, BC1D_WALL = 1, 2, 3, 4


class Params1D(ctypes.Structure):
    """Mirror of ``fvwb1d_params`` (include/fvwb.h)."""

    _fields_ = [
        ("nvar", ctypes.c_int32), ("nx", ctypes.c_int32), ("bc", ctypes.c_int32),
        ("recon", ctypes.c_int32), ("integrator", ctypes.c_int32), ("nequilibrium", ctypes.c_int32),
        ("gamma", ctypes.c_double), ("boxlen", ctypes.c_double), ("cfl", ctypes.c_double),
        ("g", ctypes.c_double),
    ]


def make_params_1d(nx, *, bc=BC1D_ZERO_GRADIENT, recon=RECON_FIRST_ORDER, integrator=RK2_HEUN,
                   nequilibrium=2, gamma=GAMMA_REF, boxlen=1.0, cfl=0.5, g=1.0) -> Params1D:
    return Params1D(3, nx, bc, recon, integrator, nequilibrium, gamma, boxlen, cfl, g)


DG1D_RIEMANN_LLF, DG1D_RIEMANN_HLLC = 1, 2


class DG1DParams(ctypes.Structure):
    """Mirror of ``fvwb_dg1d_params`` (include/fvwb.h)."""

    _fields_ = [
        ("nvar", ctypes.c_int32), ("nx", ctypes.c_int32), ("n", ctypes.c_int32),
        ("riemann", ctypes.c_int32), ("source", ctypes.c_int32), ("pad", ctypes.c_int32),
        ("gamma", ctypes.c_double), ("boxlen", ctypes.c_double),
    ]


def make_dg1d_params(nx=128, *, n=3, riemann=DG1D_RIEMANN_HLLC, source=2, gamma=GAMMA_REF,
                     boxlen=1.0) -> DG1DParams:
    """dg_commons.f90 defaults: n = 3, HLLC (riemann = 2), source 2, gamma 1.4 as FP32."""
    return DG1DParams(3, nx, n, riemann, source, 0, gamma, boxlen)


DG_LIMITER_NONE, DG_LIMITER_TVB, DG_LIMITER_POSITIVITY, DG_LIMITER_TVB_POSITIVITY = 0, 1, 2, 3
DG_LIMITER_CHARACTERISTIC, DG_LIMITER_CHARACTERISTIC_POSITIVITY = 4, 5
DG_SSPRK3, DG_SSPRK54 = 0, 1


class DGParams(ctypes.Structure):
    """Mirror of ``fvwb_dg_params`` (include/fvwb.h)."""

    _fields_ = [
        ("nvar", ctypes.c_int32), ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("degree", ctypes.c_int32), ("flux", ctypes.c_int32), ("limiter", ctypes.c_int32),
        ("bc", ctypes.c_int32), ("integrator", ctypes.c_int32),
        ("gamma", ctypes.c_double), ("boxlen_x", ctypes.c_double), ("boxlen_y", ctypes.c_double),
        ("cfl", ctypes.c_double), ("rho0", ctypes.c_double), ("p0", ctypes.c_double),
        ("g", ctypes.c_double), ("tvb_m", ctypes.c_double),
    ]


def make_dg_params(nx, ny, *, degree=2, flux=FLUX_HLL, limiter=DG_LIMITER_TVB, gamma=1.4, boxlen_x=1.0,
                   boxlen_y=1.0, cfl=0.8, rho0=1.0, p0=1.0, g=1.0, tvb_m=50.0, integrator=0) -> DGParams:
    return DGParams(4, nx, ny, degree, flux, limiter, BC_REFLECT, integrator, gamma, boxlen_x, boxlen_y, cfl,
                    rho0, p0, g, tvb_m)


def make_params(nx, ny, *, bc=BC_FROZEN_RING, recon=RECON_FIRST_ORDER, flux=FLUX_RUSANOV,
                integrator=RK2_HEUN, precision=FP64, gamma=GAMMA_REF, boxlen_x=1.0,
                boxlen_y=1.0, cfl=0.5, rho0=RHO0_REF, p0=1.0, g=1.0) -> Params:
    return Params(4, nx, ny, bc, recon, flux, integrator, precision, gamma, boxlen_x,
                  boxlen_y, cfl, rho0, p0, g)


class RDParams(ctypes.Structure):
    """orc_rd_params: the values benchmark_rd.f90 reads from its module
    parameters_rd (bc 2 periodic copies, 3 transmissive copies, 0 none)."""
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("bc", ctypes.c_int), ("pad", ctypes.c_int),
                ("gamma", ctypes.c_double), ("boxlen_x", ctypes.c_double), ("boxlen_y", ctypes.c_double),
                ("cfl", ctypes.c_double)]


RD_BC_NONE, RD_BC_PERIODIC, RD_BC_TRANSMISSIVE = 0, 2, 3


class DGQParams(ctypes.Structure):
    """Mirror of ``fvwb_dgq_params`` (include/fvwb.h): the compile-time switches
    of 2d/parameters_dg_2d.f90 for the reference's own 2D DG program."""
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("order", ctypes.c_int32),
                ("flux_type", ctypes.c_int32), ("source", ctypes.c_int32), ("grad_phi_case", ctypes.c_int32),
                ("bc", ctypes.c_int32), ("ninit", ctypes.c_int32), ("solver", ctypes.c_int32),
                ("use_limiter", ctypes.c_int32), ("gamma", ctypes.c_double), ("boxlen_x", ctypes.c_double),
                ("boxlen_y", ctypes.c_double), ("cfl", ctypes.c_double), ("eps", ctypes.c_double)]


DGQ_FLUX_NONE, DGQ_FLUX_LLF1, DGQ_FLUX_HLL2, DGQ_FLUX_HLLC = 0, 1, 2, 3
DGQ_SOLVER_RK4, DGQ_SOLVER_SS4, DGQ_SOLVER_EQL, DGQ_SOLVER_DEB = 0, 1, 2, 3
EPS_REF_DG = float(np.float32(1e-10))    # 2d/parameters_dg_2d.f90:30, default-real literal


def make_dgq_params(n=8, *, order=2, flux_type=DGQ_FLUX_NONE, source=1, grad_phi_case=2, bc=1, ninit=1,
                    solver=DGQ_SOLVER_RK4, use_limiter=1, gamma=GAMMA_REF, boxlen_x=1.0, boxlen_y=1.0,
                    cfl=float(np.float32(0.2)), eps=EPS_REF_DG) -> DGQParams:
    """Defaults: the shipped 2d/parameters_dg_2d.f90 (8x8, mx = my = 2, 'llf',
    source 1, grad_phi_case 2, bc 1, ninit 1, 'RK4', 'ONP', cfl 0.2 as a default-real
    literal)."""
    return DGQParams(n, n, order, flux_type, source, grad_phi_case, bc, ninit, solver, use_limiter, gamma,
                     boxlen_x, boxlen_y, cfl, eps)


def dgq_coords(p: DGQParams):
    """get_coords (2d/benchmark_2d_dg.f90:93-120): x(i,j,a,b) = (i-0.5) dx +
    dx/2 x_quad(a), with (i-0.5) in default REAL as the reference writes it."""
    nx, ny, o = p.nx, p.ny, p.order
    xq, _ = dg_gl_quadrature(o)
    dx, dy = p.boxlen_x / nx, p.boxlen_y / ny
    ic = (np.arange(1, nx + 1, dtype=np.float32) - np.float32(0.5)).astype(np.float64)
    jc = (np.arange(1, ny + 1, dtype=np.float32) - np.float32(0.5)).astype(np.float64)
    x = np.zeros((nx, ny, o, o), order="F")
    y = np.zeros((nx, ny, o, o), order="F")
    for a in range(o):
        for b in range(o):
            x[:, :, a, b] = (ic * dx + dx / 2.0 * xq[a])[:, None]
            y[:, :, a, b] = (jc * dy + dy / 2.0 * xq[b])[None, :]
    return x, y


def refdg_update(p: DGQParams, delta_u, x, y):
    """compute_update (2d/benchmark_2d_dg.f90:1137-1479) as shipped."""
    assert lib().orc_refdg_valid(ctypes.byref(p))
    du, x, y = fortran(delta_u), fortran(x), fortran(y)
    out = np.zeros_like(du, order="F")
    lib().orc_refdg_update(ctypes.byref(p), _p(du), _p(x), _p(y), _p(out))
    return out


def refdg_max_speed(p: DGQParams, u):
    """compute_max_speed (:826-870) on u(1:4,1:nx,1:ny): (cs_max, v_xmax, v_ymax, speed_max)."""
    out = np.zeros(4)
    lib().orc_refdg_max_speed(ctypes.byref(p), _p(fortran(u)), _p(out))
    return tuple(out)


def refdg_limiter(p: DGQParams, u):
    u = fortran(u).copy(order="F")
    lib().orc_refdg_limiter(ctypes.byref(p), _p(u))
    return u


def refdg_evolve(p: DGQParams, delta_u, x, y, t_end, max_steps=1 << 30, t0=0.0):
    """evolve's loop (:666-769) from a limited state: (modes, t, steps, dts)."""
    du = fortran(delta_u).copy(order="F")
    t = ctypes.c_double(t0)
    nmax = int(min(max_steps, 1 << 20))
    dts = np.zeros(nmax)
    n = lib().orc_refdg_evolve(ctypes.byref(p), _p(du), _p(fortran(x)), _p(fortran(y)), float(t_end), nmax,
                               ctypes.byref(t), _p(dts))
    return du, t.value, n, dts[:n]


def refdg_error(p: DGQParams, u, u_init):
    """compute_error (:23-89): (linf[4], l1[4], l2[4])."""
    out = np.zeros(12)
    lib().orc_refdg_error(p.nx, p.ny, p.order, _p(fortran(u)), _p(fortran(u_init)), _p(out))
    return out[:4], out[4:8], out[8:]


def make_rd_params(nx, ny, *, bc=RD_BC_TRANSMISSIVE, gamma=GAMMA_REF, boxlen_x=1.0, boxlen_y=1.0,
                   cfl=0.5) -> RDParams:
    return RDParams(nx, ny, bc, 0, gamma, boxlen_x, boxlen_y, cfl)


def build() -> None:
    subprocess.run(["make", "-s", "-B", "-C", HERE, "liboracle_fvwb.so"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        pp = ctypes.POINTER(Params)
        L.orc_ref2d_update_exact.argtypes = [pp, ctypes.c_int, dp, dp, dp]
        L.orc_ref2d_update_plain.argtypes = [pp, dp, dp]
        L.orc_ref2d_max_speed.argtypes = [pp, dp]
        L.orc_ref2d_max_speed.restype = ctypes.c_double
        L.orc_ref2d_evolve.argtypes = [pp, ctypes.c_int, dp, dp, ctypes.c_double, ctypes.c_int,
                                       dp, ctypes.POINTER(ctypes.c_int)]
        L.orc_fv2d_residual.argtypes = [pp, dp, dp, dp]
        L.orc_fv2d_max_speed.argtypes = [pp, dp]
        L.orc_fv2d_max_speed.restype = ctypes.c_double
        L.orc_fv2d_advance.argtypes = [pp, dp, dp, ctypes.c_double, ctypes.c_int, dp,
                                       ctypes.POINTER(ctypes.c_int), dp]
        L.orc_fv2d_face_flux.argtypes = [pp, ctypes.c_int, dp, dp, dp, dp]
        L.orc_fv2d_step.argtypes = [pp, dp, dp, ctypes.c_double]
        L.orc_minmod.argtypes = [ctypes.c_double] * 3
        L.orc_minmod.restype = ctypes.c_double
        L.orc_generalized_minmod.argtypes = [ctypes.c_double] * 5
        L.orc_generalized_minmod.restype = ctypes.c_double
        for name in ("orc_ref_llflux", "orc_ref_hllflux"):
            getattr(L, name).argtypes = [ctypes.c_double, dp, dp, dp, dp, dp]
        L.orc_ref_riemann_hllc_1d.argtypes = [ctypes.c_double, dp, dp, dp]
        L.orc_hllc_2d.argtypes = [ctypes.c_double, ctypes.c_int, dp, dp, dp]
        L.orc_ref_flux.argtypes = [ctypes.c_double, dp, dp, dp]
        L.orc_ref_speed.argtypes = [ctypes.c_double, dp]
        L.orc_ref_speed.restype = ctypes.c_double
        p1 = ctypes.POINTER(Params1D)
        L.orc_ref1d_update.argtypes = [p1, dp, dp, dp]
        L.orc_ref1d_update_sr.argtypes = [p1, dp, dp]
        L.orc_ref1d_update_fvm.argtypes = [p1, dp, dp]
        L.orc_fvm_update.argtypes = [p1, ctypes.c_int, dp, dp]
        L.orc_ref1d_max_speed.argtypes = [p1, dp]
        L.orc_ref1d_max_speed.restype = ctypes.c_double
        L.orc_ref1d_equilibrium.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int, dp, dp]
        L.orc_ref1d_llflux.argtypes = [ctypes.c_double, dp, dp, dp, dp, dp]
        L.orc_ref1d_flux.argtypes = [ctypes.c_double, dp, dp]
        L.orc_fv1d_residual.argtypes = [p1, dp, dp, dp]
        L.orc_fv1d_face_flux.argtypes = [p1, dp, dp, dp, dp]
        L.orc_fv1d_equilibrium.argtypes = [p1, ctypes.c_double, dp]
        L.orc_fv1d_advance.argtypes = [p1, dp, dp, ctypes.c_double, ctypes.c_int, dp,
                                       ctypes.POINTER(ctypes.c_int), dp]
        pd = ctypes.POINTER(DGParams)
        L.orc_dg_legendre.argtypes = [ctypes.c_double, ctypes.c_int]
        L.orc_dg_legendre.restype = ctypes.c_double
        L.orc_dg_legendre_prime.argtypes = [ctypes.c_double, ctypes.c_int]
        L.orc_dg_legendre_prime.restype = ctypes.c_double
        L.orc_dg_gl_quadrature.argtypes = [ctypes.c_int, dp, dp]
        L.orc_euler_eigenvectors.argtypes = [ctypes.c_double, ctypes.c_int, dp, dp, dp]
        L.orc_ref_dg2d_solve_for_t.argtypes = [ctypes.c_double, ctypes.c_double, dp, dp]
        L.orc_ref_dg2d_solve_for_t.restype = ctypes.c_double
        L.orc_ref_dg2d_positivity.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                              ctypes.c_double, dp]
        L.orc_dg1d_legendre.argtypes = [ctypes.c_double, ctypes.c_int]
        L.orc_dg1d_legendre.restype = ctypes.c_double
        L.orc_dg1d_legendre_prime.argtypes = [ctypes.c_double, ctypes.c_int]
        L.orc_dg1d_legendre_prime.restype = ctypes.c_double
        L.orc_dg1d_gl_quadrature.argtypes = [ctypes.c_int, dp, dp]
        L.orc_dg1d_tables.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp]
        L.orc_dg1d_update_exact_delta.argtypes = [ctypes.POINTER(DG1DParams), dp, dp, dp]
        L.orc_dg1d_rki_coeffs.argtypes = [dp]
        L.orc_dg1d_initial.argtypes = [ctypes.POINTER(DG1DParams), ctypes.c_int, ctypes.c_double, dp, dp, dp]
        L.orc_dg1d_run_rki.argtypes = [ctypes.POINTER(DG1DParams), dp, dp, dp, ctypes.c_double, ctypes.c_int, dp,
                                       dp, dp]
        L.orc_dg1d_run_rki.restype = ctypes.c_int
        L.orc_ref_dg2d_modes_from_nodes.argtypes = [ctypes.c_int] * 3 + [dp, dp]
        L.orc_ref_dg2d_nodes_from_modes.argtypes = [ctypes.c_int] * 3 + [dp, dp]
        L.orc_dg2d_nmodes.argtypes = [ctypes.c_int]
        L.orc_dg2d_residual.argtypes = [pd, dp, dp]
        L.orc_dg2d_limit.argtypes = [pd, dp]
        L.orc_dg2d_max_rate.argtypes = [pd, dp]
        L.orc_dg2d_max_rate.restype = ctypes.c_double
        L.orc_dg2d_advance.argtypes = [pd, dp, ctypes.c_double, ctypes.c_int, dp, ctypes.POINTER(ctypes.c_int), dp]
        L.orc_dg2d_ssp54_coeffs.argtypes = [dp]
        L.orc_dg2d_nodal.argtypes = [pd, dp, dp]
        L.orc_dg2d_project.argtypes = [pd, dp, dp]
        pr = ctypes.POINTER(RDParams)
        L.orc_rd_update.argtypes = [pr, dp, dp, ctypes.c_double, dp]
        L.orc_rd_update.restype = ctypes.c_int
        L.orc_rd_cell.argtypes = [pr, dp, dp, ctypes.c_double, ctypes.c_int, ctypes.c_int, dp]
        L.orc_rd_max_speed.argtypes = [pr, dp]
        L.orc_rd_max_speed.restype = ctypes.c_double
        L.orc_rd_step.argtypes = [pr, dp, ctypes.c_double, ctypes.c_double]
        L.orc_rd_step.restype = ctypes.c_double
        L.orc_rd_run.argtypes = [pr, dp, ctypes.c_double, ctypes.c_int, dp, ctypes.POINTER(ctypes.c_int)]
        L.orc_rd_run.restype = ctypes.c_int
        L.orc_rd_initial.argtypes = [pr, ctypes.c_int, ctypes.c_double, dp]
        L.orc_rd_initial.restype = ctypes.c_int
        L.orc_rd_conservative.argtypes = [ctypes.c_double, dp, dp]
        pq = ctypes.POINTER(DGQParams)
        L.orc_refdg_valid.argtypes = [pq]
        L.orc_refdg_valid.restype = ctypes.c_int
        L.orc_refdg_update.argtypes = [pq, dp, dp, dp, dp]
        L.orc_refdg_max_speed.argtypes = [pq, dp, dp]
        L.orc_refdg_dt.argtypes = [pq, dp, ctypes.c_double, ctypes.c_double, dp]
        L.orc_refdg_dt.restype = ctypes.c_double
        L.orc_refdg_limiter.argtypes = [pq, dp]
        L.orc_refdg_evolve.argtypes = [pq, dp, dp, dp, ctypes.c_double, ctypes.c_int, dp, dp]
        L.orc_refdg_evolve.restype = ctypes.c_int
        L.orc_refdg_error.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp, dp]
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim == 1)
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def fortran(a):
    return np.asfortranarray(a, dtype=np.float64)


# ---- reference restatements ------------------------------------------------
def ref2d_update_exact(p: Params, nequilibrium, u, w_eq):
    u, w_eq = fortran(u), fortran(w_eq)
    out = np.zeros_like(u, order="F")
    lib().orc_ref2d_update_exact(ctypes.byref(p), nequilibrium, _p(u), _p(w_eq), _p(out))
    return out


def ref2d_update_plain(p: Params, u):
    """compute_update (benchmark_2d.f90:370-463), the plain non-WB residual."""
    u = fortran(u)
    out = np.zeros_like(u, order="F")
    lib().orc_ref2d_update_plain(ctypes.byref(p), _p(u), _p(out))
    return out


def ref2d_max_speed(p: Params, u):
    u = fortran(u)
    return lib().orc_ref2d_max_speed(ctypes.byref(p), _p(u))


def ref2d_evolve(p: Params, nequilibrium, u, w_eq, t_end, max_steps=1 << 30):
    u = fortran(u).copy(order="F")
    w_eq = fortran(w_eq)
    t = ctypes.c_double()
    n = ctypes.c_int()
    lib().orc_ref2d_evolve(ctypes.byref(p), nequilibrium, _p(u), _p(w_eq), t_end, max_steps,
                           ctypes.byref(t), ctypes.byref(n))
    return u, t.value, n.value


# ---- configured scheme ---------------------------------------------------------
def fv2d_residual(p: Params, u, w_eq):
    u, w_eq = fortran(u), fortran(w_eq)
    out = np.zeros_like(u, order="F")
    lib().orc_fv2d_residual(ctypes.byref(p), _p(u), _p(w_eq), _p(out))
    return out


def fv2d_advance(p: Params, u, w_eq, t_end=float("inf"), max_steps=1):
    u = fortran(u).copy(order="F")
    w_eq = fortran(w_eq)
    t, dt = ctypes.c_double(), ctypes.c_double()
    n = ctypes.c_int()
    lib().orc_fv2d_advance(ctypes.byref(p), _p(u), _p(w_eq), t_end, max_steps, ctypes.byref(t),
                           ctypes.byref(n), ctypes.byref(dt))
    return u, t.value, n.value, dt.value


def fv2d_step(p: Params, u, w_eq, dt):
    u = fortran(u).copy(order="F")
    w_eq = fortran(w_eq)
    lib().orc_fv2d_step(ctypes.byref(p), _p(u), _p(w_eq), dt)
    return u


def fv2d_face_flux(p: Params, direction, ueq, dl, dr):
    out = np.zeros(4)
    ueq, dl, dr = map(_vec, (ueq, dl, dr))
    lib().orc_fv2d_face_flux(ctypes.byref(p), direction, _p(ueq), _p(dl), _p(dr), _p(out))
    return out


def minmod(x, y, z):
    return lib().orc_minmod(x, y, z)


def generalized_minmod(x, y, z, M, dx):
    return lib().orc_generalized_minmod(x, y, z, M, dx)


def _vec(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def ref_flux(gamma, u):
    u = _vec(u)
    fx, fy = np.zeros(4), np.zeros(4)
    lib().orc_ref_flux(gamma, _p(u), _p(fx), _p(fy))
    return fx, fy


def ref_llflux(gamma, ul, ur, fl, fr):
    out = np.zeros(4)
    ul, ur, fl, fr = map(_vec, (ul, ur, fl, fr))
    lib().orc_ref_llflux(gamma, _p(ul), _p(ur), _p(fl), _p(fr), _p(out))
    return out


def ref_riemann_hllc_1d(gamma, ul, ur):
    """riemann_hllc of dg_with_source.f90:1318-1374 (1D state (rho, rho v, E))."""
    out = np.zeros(3)
    ul, ur = map(_vec, (ul, ur))
    lib().orc_ref_riemann_hllc_1d(gamma, _p(ul), _p(ur), _p(out))
    return out


def hllc_2d(gamma, direction, ul, ur):
    """The 2D HLLC of the configured schemes along x (0) or y (1), full states."""
    out = np.zeros(4)
    ul, ur = map(_vec, (ul, ur))
    lib().orc_hllc_2d(gamma, direction, _p(ul), _p(ur), _p(out))
    return out


def ref_hllflux(gamma, ul, ur, fl, fr):
    out = np.zeros(4)
    ul, ur, fl, fr = map(_vec, (ul, ur, fl, fr))
    lib().orc_ref_hllflux(gamma, _p(ul), _p(ur), _p(fl), _p(fr), _p(out))
    return out


# ---- problem set-ups shared by tests, smoke and bench ---------------------------
def equilibrium_primitive(p: Params):
    """Closed-form equilibrium at cell centres, primitive AoS(4,nx,ny)."""
    nx, ny = p.nx, p.ny
    dx, dy = p.boxlen_x / nx, p.boxlen_y / ny
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dy
    X, Y = np.meshgrid(x, y, indexing="ij")
    k = p.rho0 * p.g / p.p0
    e = np.exp(-(k * (X + Y)))
    w = np.zeros((4, nx, ny), order="F")
    w[0] = p.rho0 * e
    w[3] = p.p0 * e
    return w


def conservative(gamma, w):
    u = np.zeros_like(w, order="F")
    u[0] = w[0]
    u[1] = w[0] * w[1]
    u[2] = w[0] * w[2]
    u[3] = w[3] / (gamma - 1.0) + 0.5 * (w[0] * (w[1] * w[1] + w[2] * w[2]))
    return u


def hash_noise(nx, ny, salt):
    """Deterministic noise in [-1,1) from an integer hash of (i,j,salt);
    identical on every platform (no RNG stream involved)."""
    i = np.arange(nx, dtype=np.uint64)[:, None]
    j = np.arange(ny, dtype=np.uint64)[None, :]
    h = (i * np.uint64(73856093)) ^ (j * np.uint64(19349663)) ^ np.uint64(salt * 83492791 + 17)
    h = (h * np.uint64(0x9E3779B97F4A7C15)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    h ^= h >> np.uint64(29)
    h = (h * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    h ^= h >> np.uint64(32)
    return (h >> np.uint64(11)).astype(np.float64) / float(1 << 53) * 2.0 - 1.0


def perturbed_state(p: Params, amp_bump=1e-3, amp_noise=1e-4, x0=0.3, y0=0.3, width=100.0,
                    salt=1):
    """u = U_eq + (0,0,0,dp/(gamma-1)) for a Gaussian pressure bump plus hash noise
    of relative size amp_noise on every conservative variable."""
    w = equilibrium_primitive(p)
    ueq = conservative(p.gamma, w)
    nx, ny = p.nx, p.ny
    dx, dy = p.boxlen_x / nx, p.boxlen_y / ny
    X, Y = np.meshgrid((np.arange(nx) + 0.5) * dx, (np.arange(ny) + 0.5) * dy, indexing="ij")
    u = ueq.copy(order="F")
    u[3] += amp_bump * np.exp(-width * ((X - x0) ** 2 + (Y - y0) ** 2)) / (p.gamma - 1.0)
    if amp_noise:
        for v in range(4):
            scale = ueq[0] if v < 3 else ueq[3]
            u[v] += amp_noise * scale * hash_noise(nx, ny, salt * 4 + v)
    return u, w


# ---- 1D (benchmark_1d.f90) ---------------------------------------------------------
def ref1d_update(p: Params1D, u, w_eq):
    u, w_eq = fortran(u), fortran(w_eq)
    out = np.zeros_like(u, order="F")
    lib().orc_ref1d_update(ctypes.byref(p), _p(u), _p(w_eq), _p(out))
    return out


def ref1d_update_sr(p: Params1D, u):
    """compute_update_sr (benchmark_1d.f90:553-747), the 'WB1' residual."""
    u = fortran(u)
    out = np.zeros_like(u, order="F")
    lib().orc_ref1d_update_sr(ctypes.byref(p), _p(u), _p(out))
    return out


def ref1d_update_fvm(p: Params1D, u):
    """compute_update_fvm (benchmark_1d.f90:454-549), the plain 'FVM' residual."""
    u = fortran(u)
    out = np.zeros_like(u, order="F")
    lib().orc_ref1d_update_fvm(ctypes.byref(p), _p(u), _p(out))
    return out


def fvm_update(p: Params1D, u, source=2):
    """compute_update of fvm.f90 (:188-251); bc from p (1 or 2), source 1 or 2."""
    u = fortran(u)
    out = np.zeros_like(u, order="F")
    lib().orc_fvm_update(ctypes.byref(p), int(source), _p(u), _p(out))
    return out


def ref1d_max_speed(p: Params1D, u):
    return lib().orc_ref1d_max_speed(ctypes.byref(p), _p(fortran(u)))


def ref1d_flux(gamma, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    f = np.zeros(3)
    lib().orc_ref1d_flux(gamma, _p(u), _p(f))
    return f


def ref1d_llflux(gamma, ul, ur, fl, fr):
    ul, ur, fl, fr = (np.ascontiguousarray(a, dtype=np.float64) for a in (ul, ur, fl, fr))
    out = np.zeros(3)
    lib().orc_ref1d_llflux(gamma, _p(ul), _p(ur), _p(fl), _p(fr), _p(out))
    return out


def fv1d_residual(p: Params1D, u, w_eq):
    u, w_eq = fortran(u), fortran(w_eq)
    out = np.zeros_like(u, order="F")
    lib().orc_fv1d_residual(ctypes.byref(p), _p(u), _p(w_eq), _p(out))
    return out


def fv1d_face_flux(p: Params1D, ueq, dl, dr):
    ueq, dl, dr = (np.ascontiguousarray(a, dtype=np.float64) for a in (ueq, dl, dr))
    out = np.zeros(3)
    lib().orc_fv1d_face_flux(ctypes.byref(p), _p(ueq), _p(dl), _p(dr), _p(out))
    return out


def fv1d_advance(p: Params1D, u, w_eq, t_end=float("inf"), max_steps=1 << 30):
    u = fortran(u).copy(order="F")
    w_eq = fortran(w_eq)
    t, dt = ctypes.c_double(), ctypes.c_double()
    n = ctypes.c_int()
    lib().orc_fv1d_advance(ctypes.byref(p), _p(u), _p(w_eq), t_end, max_steps, ctypes.byref(t),
                           ctypes.byref(n), ctypes.byref(dt))
    return u, t.value, n.value, dt.value


def equilibrium_1d(p: Params1D, x):
    """Closed-form equilibrium of p at points x, primitive AoS(3,n)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    w = np.zeros((3, len(x)), order="F")
    for i, xv in enumerate(x):
        wi = np.zeros(3)
        lib().orc_fv1d_equilibrium(ctypes.byref(p), float(xv), _p(wi))
        w[:, i] = wi
    return w


def conservative_1d(gamma, w):
    u = np.zeros_like(w, order="F")
    u[0] = w[0]
    u[1] = w[0] * w[1]
    u[2] = w[2] / (gamma - 1.0) + 0.5 * w[0] * (w[1] * w[1])
    return u


def perturbed_state_1d(p: Params1D, amp_bump=1e-3, x0=0.5, width=0.05, amp_noise=0.0, salt=1):
    """Equilibrium plus a Gaussian pressure pulse amp*exp(-(x-x0)^2/(2 width^2))
    (the fvm.f90:124 convention) and hash noise of relative size amp_noise on
    every conservative variable; returns (u, w_eq) AoS(3,nx)."""
    nx = p.nx
    x = (np.arange(nx) + 0.5) * (p.boxlen / nx)
    w = equilibrium_1d(p, x)
    u = conservative_1d(p.gamma, w)
    u[2] += amp_bump * np.exp(-((x - x0) ** 2) / (2.0 * width ** 2)) / (p.gamma - 1.0)
    if amp_noise:
        for v in range(3):
            scale = u[0] if v < 2 else u[2]
            u[v] += amp_noise * scale * hash_noise(nx, 1, salt * 3 + v)[:, 0]
    return u, w


# ---- 2D modal DG ---------------------------------------------------------------------
def dg_legendre(x, n):
    return lib().orc_dg_legendre(float(x), int(n))


def dg_legendre_prime(x, n):
    return lib().orc_dg_legendre_prime(float(x), int(n))


def dg_gl_quadrature(n):
    x, w = np.zeros(n), np.zeros(n)
    lib().orc_dg_gl_quadrature(n, _p(x), _p(w))
    return x, w


def euler_eigenvectors(gamma, direction, u):
    """(R, L): right eigenvectors as columns and their inverse, of the x (0) or y (1)
    Euler flux Jacobian at the conservative state u."""
    u = np.ascontiguousarray(u, dtype=float)
    R, L = np.zeros(16), np.zeros(16)
    lib().orc_euler_eigenvectors(float(gamma), int(direction), _p(u), _p(R), _p(L))
    return R.reshape(4, 4), L.reshape(4, 4)


def ref_dg2d_solve_for_t(gamma, eps, u, u_avg):
    u, ua = np.ascontiguousarray(u, dtype=float), np.ascontiguousarray(u_avg, dtype=float)
    return lib().orc_ref_dg2d_solve_for_t(float(gamma), float(eps), _p(u), _p(ua))


def ref_dg2d_positivity(u, gamma=GAMMA_REF, eps=float(np.float32(1e-10))):
    """compute_positivity (2d/limiters.f90:478-654) on tensor modes AoS(4,nx,ny,o,o)."""
    u = fortran(u).copy(order="F")
    _, nx, ny, o, _ = u.shape
    lib().orc_ref_dg2d_positivity(nx, ny, o, float(gamma), float(eps), _p(u))
    return u


# ---- 1D modal DG (dg_with_source.f90, legendre.f90) -----------------------------------
def dg1d_legendre(x, n):
    return lib().orc_dg1d_legendre(float(x), int(n))


def dg1d_legendre_prime(x, n):
    return lib().orc_dg1d_legendre_prime(float(x), int(n))


def dg1d_gl_quadrature(n):
    x, w = np.zeros(n), np.zeros(n)
    lib().orc_dg1d_gl_quadrature(n, _p(x), _p(w))
    return x, w


def dg1d_tables(n):
    """(chsi, w, phi[i,j] = phi_i(chsi_j), dphi[i,j], phi_lr[i] = (phi_i(-1), phi_i(+1)))."""
    c, w, ph, dph, lr = np.zeros(n), np.zeros(n), np.zeros(n * n), np.zeros(n * n), np.zeros(2 * n)
    lib().orc_dg1d_tables(n, _p(c), _p(w), _p(ph), _p(dph), _p(lr))
    return c, w, ph.reshape(n, n), dph.reshape(n, n), lr.reshape(n, 2)


def dg1d_equilibrium_nodes(p: DG1DParams):
    """u_eq(1:3,1:n,1:nx) as the reference's main program builds it
    (dg_with_source.f90:71-89): exp(-x) at x = xcell + dx/2*chsi_j, conservative."""
    chsi, _ = dg1d_gl_quadrature(p.n)
    dx = p.boxlen / p.nx
    xc = (np.arange(1, p.nx + 1) - 0.5) * dx
    x = xc[None, :] + dx / 2.0 * chsi[:, None]
    r = np.exp(-x)
    u = np.zeros((3, p.n, p.nx), order="F")
    u[0] = r
    u[2] = r / (p.gamma - 1.0)
    return u


def dg1d_update_exact_delta(p: DG1DParams, delta_u, u_eq):
    delta_u, u_eq = fortran(delta_u), fortran(u_eq)
    out = np.zeros_like(delta_u, order="F")
    lib().orc_dg1d_update_exact_delta(ctypes.byref(p), _p(delta_u), _p(u_eq), _p(out))
    return out


def dg1d_rki_coeffs():
    c = np.zeros(15)
    lib().orc_dg1d_rki_coeffs(_p(c))
    return c


def dg1d_initial(p: DG1DParams, ninit=8, pert=float(np.float32(1e-8))):
    """(delta_u, u_eq, u_init) of the reference main program's set-up, AoS(3,n,nx)."""
    shape = (3, p.n, p.nx)
    d, e, u = (np.zeros(shape, order="F") for _ in range(3))
    lib().orc_dg1d_initial(ctypes.byref(p), int(ninit), float(pert), _p(d), _p(e), _p(u))
    return d, e, u


def dg1d_run_rki(p: DG1DParams, delta_u, u_eq, u_init, t_end, max_steps=1 << 30):
    """The main loop with 'RKi'; returns (delta_u, u_init, t, steps, ts, dts)."""
    d, u = fortran(delta_u).copy(order="F"), fortran(u_init).copy(order="F")
    e = fortran(u_eq)
    cap = int(min(max_steps, 1 << 20))
    ts, dts, t = np.zeros(cap), np.zeros(cap), ctypes.c_double()
    k = lib().orc_dg1d_run_rki(ctypes.byref(p), _p(d), _p(e), _p(u), float(t_end), cap, ctypes.byref(t),
                               _p(ts), _p(dts))
    return d, u, t.value, k, ts[:k], dts[:k]


def ref_dg2d_modes_from_nodes(nodes):
    nodes = fortran(nodes)
    _, nx, ny, o, _ = nodes.shape
    out = np.zeros_like(nodes, order="F")
    lib().orc_ref_dg2d_modes_from_nodes(nx, ny, o, _p(nodes), _p(out))
    return out


def ref_dg2d_nodes_from_modes(modes):
    modes = fortran(modes)
    _, nx, ny, o, _ = modes.shape
    out = np.zeros_like(modes, order="F")
    lib().orc_ref_dg2d_nodes_from_modes(nx, ny, o, _p(modes), _p(out))
    return out


def dg_nmodes(degree):
    return (degree + 1) * (degree + 2) // 2


def dg2d_residual(p: DGParams, d):
    d = fortran(d)
    out = np.zeros_like(d, order="F")
    lib().orc_dg2d_residual(ctypes.byref(p), _p(d), _p(out))
    return out


def dg2d_limit(p: DGParams, d):
    d = fortran(d).copy(order="F")
    lib().orc_dg2d_limit(ctypes.byref(p), _p(d))
    return d


def dg2d_max_rate(p: DGParams, d):
    return lib().orc_dg2d_max_rate(ctypes.byref(p), _p(fortran(d)))


def dg2d_advance(p: DGParams, d, t_end=float("inf"), max_steps=1 << 30):
    d = fortran(d).copy(order="F")
    t, dt = ctypes.c_double(), ctypes.c_double()
    n = ctypes.c_int()
    lib().orc_dg2d_advance(ctypes.byref(p), _p(d), t_end, max_steps, ctypes.byref(t), ctypes.byref(n),
                           ctypes.byref(dt))
    return d, t.value, n.value, dt.value


def dg2d_ssp54_coeffs():
    """Coefficients of the reference's SSP(5,4) (2d/benchmark_2d_dg.f90:686-707) as
    the oracle uses them: c10, a20, a21, c21, a30, a32, c32, a40, a43, c43, a50, a52,
    a53, a54, c53, c54."""
    c = np.zeros(16)
    lib().orc_dg2d_ssp54_coeffs(_p(c))
    return c


def dg2d_nodal(p: DGParams, d):
    d = fortran(d)
    nq = p.degree + 1
    out = np.zeros((4, p.nx, p.ny, nq * nq), order="F")
    lib().orc_dg2d_nodal(ctypes.byref(p), _p(d), _p(out))
    return out


def dg2d_project(p: DGParams, nodal):
    nodal = fortran(nodal)
    out = np.zeros((4, p.nx, p.ny, dg_nmodes(p.degree)), order="F")
    lib().orc_dg2d_project(ctypes.byref(p), _p(nodal), _p(out))
    return out


def dg2d_points(p: DGParams):
    """Physical coordinates (X, Y) of the volume points, shape (nx, ny, nq*nq)."""
    nq = p.degree + 1
    xq, _ = dg_gl_quadrature(nq)
    dx, dy = p.boxlen_x / p.nx, p.boxlen_y / p.ny
    xc = (np.arange(p.nx) + 0.5) * dx
    yc = (np.arange(p.ny) + 0.5) * dy
    X = np.zeros((p.nx, p.ny, nq * nq))
    Y = np.zeros((p.nx, p.ny, nq * nq))
    for b in range(nq):
        for a in range(nq):
            X[:, :, a + nq * b] = (xc + 0.5 * dx * xq[a])[:, None]
            Y[:, :, a + nq * b] = (yc + 0.5 * dy * xq[b])[None, :]
    return X, Y


def fourier_modes(n=16, seed=0xF00D, mmax=8):
    """(a_k, m_k, n_k, theta_k) of the DG perturbation: a ~ U[0.5,1], (m,n) in
    {1..mmax}^2, theta ~ U[0, 2 pi), from numpy's PCG64 seeded with `seed`."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.5, 1.0, n)
    mk = rng.integers(1, mmax + 1, n)
    nk = rng.integers(1, mmax + 1, n)
    th = rng.uniform(0.0, 2 * np.pi, n)
    return a, mk, nk, th


def dg2d_fourier_state(p: DGParams, amp=1e-4, modes=None):
    """Deviation modes of a pressure perturbation
    dp = amp * sum_k a_k cos(2 pi (m_k x + n_k y) + theta_k) / sum_k |a_k|
    (dE = dp/(gamma-1), other deviations 0), projected on P_k."""
    a, mk, nk, th = fourier_modes() if modes is None else modes
    X, Y = dg2d_points(p)
    dp = np.zeros_like(X)
    for q in range(len(a)):
        dp += a[q] * np.cos(2 * np.pi * (mk[q] * X + nk[q] * Y) + th[q])
    dp *= amp / np.sum(np.abs(a))
    nodal = np.zeros((4, p.nx, p.ny, X.shape[2]), order="F")
    nodal[3] = dp / (p.gamma - 1.0)
    return dg2d_project(p, nodal)


# ---- residual distribution (benchmark_rd.f90) ----------------------------------
def rd_update(p: RDParams, u, diff, dt):
    """compute_update_rd: dudt AoS(4,nx,ny) of the state u and the correction diff."""
    u, diff = fortran(u), fortran(diff)
    out = np.zeros((4, p.nx, p.ny), order="F")
    rc = lib().orc_rd_update(ctypes.byref(p), _p(u), _p(diff), float(dt), _p(out))
    if rc != 0:
        raise ValueError("orc_rd_update: bad parameters")
    return out


def rd_cell(p: RDParams, u, diff, dt, i, j):
    """Corner residuals (11, 12, 21, 22) x 4 variables of cell (i, j)."""
    u, diff = fortran(u), fortran(diff)
    phi = np.zeros(16)
    lib().orc_rd_cell(ctypes.byref(p), _p(u), _p(diff), float(dt), int(i), int(j), _p(phi))
    return phi.reshape(4, 4)


def rd_max_speed(p: RDParams, u):
    return lib().orc_rd_max_speed(ctypes.byref(p), _p(fortran(u)))


def rd_step(p: RDParams, u, t=0.0, t_end=float("inf")):
    """One 'RD1' step in place; returns dt."""
    assert u.flags.f_contiguous and u.dtype == np.float64
    dt = lib().orc_rd_step(ctypes.byref(p), _p(u), float(t), float(t_end))
    if dt < 0:
        raise ValueError("orc_rd_step: bad parameters")
    return dt


def rd_run(p: RDParams, u, t_end, max_steps=1 << 30):
    """evolve's loop; returns (u, t, steps)."""
    u = fortran(u).copy(order="F")
    t = ctypes.c_double(0.0)
    n = ctypes.c_int(0)
    if lib().orc_rd_run(ctypes.byref(p), _p(u), float(t_end), int(max_steps), ctypes.byref(t), ctypes.byref(n)):
        raise ValueError("orc_rd_run: bad parameters")
    return u, t.value, n.value


def rd_initial(p: RDParams, ninit, eta=1e-3):
    u = np.zeros((4, p.nx, p.ny), order="F")
    if lib().orc_rd_initial(ctypes.byref(p), int(ninit), float(eta), _p(u)):
        raise ValueError(f"orc_rd_initial: no case {ninit}")
    return u


def rd_conservative(gamma, w):
    w = np.ascontiguousarray(w, dtype=np.float64)
    u = np.zeros(4)
    lib().orc_rd_conservative(float(gamma), _p(w), _p(u))
    return u
