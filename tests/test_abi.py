"""The C ABI library loads and exports every symbol include/fvwb.h declares,
and argument errors are reported without touching the GPU."""
import ctypes

import pytest

import fvwb


def test_library_exports_every_declared_symbol():
    L = fvwb.lib()
    declared = fvwb.declared_symbols()
    assert len(declared) >= 14
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing


def test_reference_params_match_parameters_2d():
    p = fvwb.reference_params(300, 300, 2)
    assert (p.nvar, p.nx, p.ny) == (4, 300, 300)
    assert p.gamma == 1.399999976158142 and p.rho0 == 1.2100000381469727
    assert (p.bc, p.recon, p.flux, p.integrator, p.precision) == (
        fvwb.BC_FROZEN_RING, fvwb.RECON_FIRST_ORDER, fvwb.FLUX_RUSANOV, fvwb.RK2_HEUN, fvwb.FP64)
    assert p.cfl == 0.5 and p.boxlen_x == 1.0 and p.boxlen_y == 1.0
    q = fvwb.reference_params(64, 64, 1)
    assert q.rho0 == 1.0 and q.p0 == 1.0 and q.g == 1.0
    with pytest.raises(fvwb.FvwbError):
        fvwb.reference_params(8, 8, 4)


@pytest.mark.parametrize("field,value", [("nvar", 3), ("nx", 2), ("ny", 3), ("bc", 7), ("recon", 5),
                                         ("flux", 9), ("integrator", 4), ("precision", 3),
                                         ("gamma", 1.0), ("cfl", 0.0), ("rho0", -1.0)])
def test_create_rejects_bad_params_before_any_device_call(field, value):
    p = fvwb.make_params(16, 16)
    setattr(p, field, value)
    h = ctypes.c_void_p()
    rc = fvwb.lib().fvwb_create(ctypes.byref(h), ctypes.byref(p))
    assert rc == fvwb.ERR_ARG and not h.value
    assert fvwb.lib().fvwb_last_error(None)


def test_null_handle_calls_are_argument_errors():
    L = fvwb.lib()
    assert L.fvwb_advance(None, 1, 1.0, None, None) == fvwb.ERR_ARG
    assert L.fvwb_download(None, None) == fvwb.ERR_ARG
    assert L.fvwb_compute_update_exact(None, None, None, None) == fvwb.ERR_ARG
    assert L.fvwb_destroy(None) == fvwb.OK


@pytest.mark.parametrize("field,value", [("nvar", 4), ("nx", 2), ("n", 0), ("n", 7), ("riemann", 0),
                                         ("riemann", 3), ("source", 3), ("gamma", 1.0), ("boxlen", 0.0)])
def test_dg1d_create_rejects_bad_params_before_any_device_call(field, value):
    p = fvwb.reference_params_dg1d(32)
    assert (p.nvar, p.nx, p.n, p.riemann, p.source) == (3, 32, 3, fvwb.DG1D_RIEMANN_HLLC, 2)
    assert p.gamma == 1.399999976158142 and p.boxlen == 1.0
    setattr(p, field, value)
    h = ctypes.c_void_p()
    rc = fvwb.lib().fvwb_dg1d_create(ctypes.byref(h), ctypes.byref(p))
    assert rc == fvwb.ERR_ARG and not h.value
    assert fvwb.lib().fvwb_dg1d_last_error(None)


def test_dg1d_null_handle_calls_are_argument_errors():
    L = fvwb.lib()
    assert L.fvwb_dg1d_compute_update_exact_delta(None, None, None, None) == fvwb.ERR_ARG
    assert L.fvwb_dg1d_destroy(None) == fvwb.OK


def test_checkpoint_null_arguments_are_argument_errors():
    L = fvwb.lib()
    for name in ("fvwb_checkpoint_save", "fvwb_checkpoint_load", "fvwb1d_checkpoint_save", "fvwb1d_checkpoint_load",
                 "fvwb_dg_checkpoint_save", "fvwb_dg_checkpoint_load", "fvwb_dg1d_checkpoint_save",
                 "fvwb_dg1d_checkpoint_load"):
        assert getattr(L, name)(None, b"/tmp/x") == fvwb.ERR_ARG, name


def test_dgq_reference_params_match_parameters_dg_2d():
    p = fvwb.reference_params_dgq()
    assert (p.nx, p.ny, p.order, p.flux_type, p.source, p.grad_phi_case, p.bc, p.ninit, p.solver,
            p.use_limiter) == (8, 8, 2, fvwb.DGQ_FLUX_NONE, 1, 2, 1, 1, fvwb.DGQ_SOLVER_RK4, 1)
    assert p.gamma == 1.399999976158142 and p.cfl == 0.20000000298023224 and p.eps == 1.0000000133514320e-10
    assert p.boxlen_x == 1.0 and p.boxlen_y == 1.0


@pytest.mark.parametrize("field,value", [("ny", 9), ("nx", 0), ("order", 1), ("order", 4), ("flux_type", 4),
                                         ("source", 0), ("source", 4), ("grad_phi_case", 3), ("bc", 0),
                                         ("solver", 4), ("gamma", 1.0), ("cfl", 0.0), ("boxlen_y", -1.0)])
def test_dgq_create_rejects_bad_params_before_any_device_call(field, value):
    p = fvwb.reference_params_dgq()
    setattr(p, field, value)
    h = ctypes.c_void_p()
    rc = fvwb.lib().fvwb_dgq_create(ctypes.byref(h), ctypes.byref(p))
    assert rc == fvwb.ERR_ARG and not h.value
    assert fvwb.lib().fvwb_dgq_last_error(None)


def test_dgq_null_handle_calls_are_argument_errors():
    L = fvwb.lib()
    assert L.fvwb_dgq_compute_update(None, None, None, None, None, None) == fvwb.ERR_ARG
    assert L.fvwb_dgq_advance(None, 1, 1.0) == fvwb.ERR_ARG
    assert L.fvwb_dgq_compute_error(None, None, None, None, None, None) == fvwb.ERR_ARG
    assert L.fvwb_dgq_destroy(None) == fvwb.OK


def test_dgq_gl_quadrature_equals_reference_build():
    """The library's host Gauss-Legendre rule (the nodes get_coords and every
    reference-DG drop-in use) against gl_quadrature of the reference build,
    n = 1..6, bit for bit (host code: no GPU)."""
    import os
    import numpy as np
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dg2d_gl_quadrature.npz")
    with np.load(gold) as z:
        for n in range(1, 7):
            x, w = fvwb.dgq_gl_quadrature(n)
            np.testing.assert_array_equal(x, z[f"x{n}"])
            np.testing.assert_array_equal(w, z[f"w{n}"])
    with pytest.raises(fvwb.FvwbError):
        fvwb.dgq_gl_quadrature(7)
