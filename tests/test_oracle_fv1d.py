"""The 1D CPU oracle: the benchmark_1d.f90 restatements ('EQL' compute_update,
'WB1' compute_update_sr, 'FVM' compute_update_fvm) and fvm.f90's compute_update
against the fixtures
written from the flang-built reference (bit for bit) and live against
oracle/_ref when present; the configured 1D scheme against what the
mathematics fixes (exact rest, flux consistency, wall symmetry and mass
conservation, convergence orders, Sod's exact solution) and against the 2D
oracle on a y-uniform problem."""
import os

import numpy as np
import pytest

import pyoracle as po
import ref_runner as rr
from make_golden import bench1d_inputs
from test_oracle_invariants import exact_riemann_density

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(GOLD, name + ".npz")) as z:
        return {k: z[k] for k in z.files}


# ---- the reference restatement -------------------------------------------------------
def test_ref1d_update_matches_reference_golden_bitwise():
    g = load("bench1d_update_128")
    p, u, w = bench1d_inputs()
    np.testing.assert_array_equal(u, g["u"])
    np.testing.assert_array_equal(w, g["w_eq"])
    np.testing.assert_array_equal(po.ref1d_update(p, u, w), g["dudt"])
    assert po.ref1d_max_speed(p, u) == float(g["cmax"][0])
    ue = po.conservative_1d(p.gamma, w)
    np.testing.assert_array_equal(po.ref1d_update(p, ue, w), g["dudt_eq"])
    assert np.count_nonzero(g["dudt_eq"]) == 0


def test_ref1d_update_polytropic_matches_reference_golden_bitwise():
    """compute_update with the polytropic equilibrium (nequilibrium 3): a flang
    build of benchmark_1d.f90 with that switch edited (oracle/Makefile); the
    state, its rest and the wave-speed maximum, bit for bit."""
    g = load("bench1d_update_neq3_128")
    p, u, w = bench1d_inputs(salt=3, nequilibrium=3)
    np.testing.assert_array_equal(u, g["u"])
    np.testing.assert_array_equal(w, g["w_eq"])
    np.testing.assert_array_equal(po.ref1d_update(p, u, w), g["dudt"])
    assert po.ref1d_max_speed(p, u) == float(g["cmax"][0])
    np.testing.assert_array_equal(po.ref1d_update(p, po.conservative_1d(p.gamma, w), w), g["dudt_eq"])
    assert np.count_nonzero(g["dudt_eq"]) == 0


def test_ref1d_update_sr_matches_reference_golden_bitwise():
    """compute_update_sr (benchmark_1d.f90:553-747, the 'WB1' solver) on the
    perturbed state, the isothermal rest state, the isentropic rest state and a
    jump with flow: bit for bit (the oracle calls the same libm pow)."""
    g = load("bench1d_update_sr_128")
    p, u, _ = bench1d_inputs()
    np.testing.assert_array_equal(u, g["u"])
    for k in ("", "_eq", "_isen", "_jump"):
        np.testing.assert_array_equal(po.ref1d_update_sr(p, g["u" + k]), g["dudt" + k], err_msg=k)


def isentropic_rest(p):
    x = (np.arange(p.nx) + 0.5) * (p.boxlen / p.nx)
    w = np.zeros((3, p.nx), order="F")
    w[0] = (1 - ((p.gamma - 1) / p.gamma) * x) ** (1 / (p.gamma - 1))
    w[2] = w[0] ** p.gamma
    return po.conservative_1d(p.gamma, w)


@pytest.mark.parametrize("nx", [16, 128, 1000])
def test_ref1d_update_sr_keeps_isentropic_rest(nx):
    """The hydrostatic reconstruction is exact for p = rho^gamma at rest under
    phi = x: every face state equals the equilibrium there, the Rusanov fluxes
    reduce to the face pressure and the pressure source cancels their
    difference, so dudt is rounding only (a few ulps of p, times nx)."""
    p = po.make_params_1d(nx)
    d = po.ref1d_update_sr(p, isentropic_rest(p))
    assert np.max(np.abs(d)) <= 64 * np.finfo(float).eps * nx


def test_ref1d_update_sr_converges_to_the_euler_gravity_operator():
    """On a smooth moving state the residual tends to -(F(U))_x + S with
    S = (0, -rho, -rho v) (phi = x), at first order (the Rusanov diffusion);
    a dropped source term or a wrong sign leaves an O(1) error instead."""
    errs = []
    for nx in (100, 200, 400):
        p = po.make_params_1d(nx)
        gm = p.gamma
        x = (np.arange(nx) + 0.5) / nx
        k = 2 * np.pi
        r, rx = 1 + 0.2 * np.sin(k * x), 0.2 * k * np.cos(k * x)
        v, vx = 0.3 * np.cos(k * x), -0.3 * k * np.sin(k * x)
        q, qx = 1 + 0.1 * np.cos(k * x), -0.1 * k * np.sin(k * x)
        w = np.asfortranarray(np.stack([r, v, q]))
        d = po.ref1d_update_sr(p, po.conservative_1d(gm, w))
        E, Ex = q / (gm - 1) + 0.5 * r * v * v, qx / (gm - 1) + 0.5 * (rx * v * v + 2 * r * v * vx)
        exact = np.stack([-(rx * v + r * vx),
                          -(rx * v * v + 2 * r * v * vx + qx) - r,
                          -((Ex + qx) * v + (E + q) * vx) - r * v])
        errs.append(np.max(np.abs(d - exact)[:, 2:-2]))
    assert errs[0] < 0.2
    assert errs[0] / errs[1] > 1.8 and errs[1] / errs[2] > 1.8, errs


def test_ref1d_update_fvm_matches_reference_golden_bitwise():
    """compute_update_fvm (benchmark_1d.f90:454-549, the plain 'FVM' solver)."""
    g = load("bench1d_update_fvm_128")
    p, u, _ = bench1d_inputs()
    np.testing.assert_array_equal(u, g["u"])
    for k in ("", "_jump"):
        np.testing.assert_array_equal(po.ref1d_update_fvm(p, g["u" + k]), g["dudt" + k], err_msg=k)


def test_fvm_update_matches_reference_golden_bitwise():
    """compute_update of fvm.f90 (:188-251) as the reference compiles it (nx=200,
    bc 2, source 2) on its own Sod set-up, a smooth moving state and condinit's
    case-7 state."""
    g = load("fvm_update_200")
    p = po.make_params_1d(200, bc=po.BC1D_ZERO_GRADIENT)
    for k in ("_sod", "_smooth", "_poly"):
        np.testing.assert_array_equal(po.fvm_update(p, g["u" + k], 2), g["dudt" + k], err_msg=k)


def test_fvm_update_periodic_is_translation_invariant():
    """bc 1 (which the reference build cannot reach): without the source every
    cell sees its wrapped neighbours the same way, so a cyclic shift of the state
    shifts dudt bit for bit; with zero gradient the end cells differ."""
    nx = 64
    u = np.asfortranarray(load("fvm_update_200")["u_smooth"][:, :nx])
    pp = po.make_params_1d(nx, bc=po.BC1D_PERIODIC)
    d = po.fvm_update(pp, u, 1)
    for k in (1, 17, 63):
        np.testing.assert_array_equal(po.fvm_update(pp, np.roll(u, k, axis=1), 1), np.roll(d, k, axis=1))
    pz = po.make_params_1d(nx, bc=po.BC1D_ZERO_GRADIENT)
    dz = po.fvm_update(pz, u, 1)
    np.testing.assert_array_equal(dz[:, 1:-1], d[:, 1:-1])
    assert not np.array_equal(dz[:, 0], d[:, 0])


def test_fvm_update_source_switch_adds_the_gravity_source():
    """source 2 minus source 1 is the closed-form source: -rho and -rho*v inside,
    half of that in the two end cells (x- and x+ clamped to the cell centre),
    to the rounding of the flux part; a uniform state at rest gives exactly zero
    with source 1 (flux consistency)."""
    nx = 50
    g = load("fvm_update_200")
    u = np.asfortranarray(g["u_smooth"][:, :nx])
    p = po.make_params_1d(nx, bc=po.BC1D_ZERO_GRADIENT)
    d1, d2 = po.fvm_update(p, u, 1), po.fvm_update(p, u, 2)
    rho, v = u[0], u[1] / u[0]
    half = np.ones(nx); half[[0, -1]] = 0.5
    tol = 8 * np.finfo(float).eps * (np.abs(d1) + np.abs(d2) + 1)
    np.testing.assert_array_equal(d2[0], d1[0])
    assert np.all(np.abs((d2[1] - d1[1]) - (-rho * half)) <= tol[1])
    assert np.all(np.abs((d2[2] - d1[2]) - (-rho * v * half)) <= tol[2])
    uni = np.asfortranarray(np.tile([[1.3], [0.0], [2.0]], (1, nx)))
    assert np.count_nonzero(po.fvm_update(p, uni, 1)) == 0


def test_ref1d_llflux_matches_reference_golden_bitwise():
    g = load("bench1d_llflux")
    gamma = float(g["gamma"])
    for q in range(len(g["f"])):
        np.testing.assert_array_equal(po.ref1d_flux(gamma, g["ul"][q]), g["fl"][q])
        np.testing.assert_array_equal(po.ref1d_llflux(gamma, g["ul"][q], g["ur"][q], g["fl"][q], g["fr"][q]),
                                      g["f"][q])


@pytest.mark.skipif(not rr.available("bench1d"), reason="oracle/_ref not built")
def test_ref1d_update_matches_live_reference_on_other_states():
    for salt in (2, 3):
        p, u, w = bench1d_inputs(salt=salt)
        u[1] *= 50.0          # stronger flow: wave speeds differ across faces
        d = rr.run("bench1d_update", u=u, w_eq=w)["dudt"]
        np.testing.assert_array_equal(po.ref1d_update(p, u, w), d)


def test_ref1d_bc_code_does_not_change_dudt():
    """benchmark_1d.f90:374-375 overwrite both boundary cells, so the bc code
    (which only changes the two boundary faces) cannot reach dudt."""
    p, u, w = bench1d_inputs()
    ds = []
    for bc in (po.BC1D_PERIODIC, po.BC1D_ZERO_GRADIENT, po.BC1D_REFLEXIVE):
        p.bc = bc
        ds.append(po.ref1d_update(p, u, w))
    np.testing.assert_array_equal(ds[0], ds[1])
    np.testing.assert_array_equal(ds[0], ds[2])
    np.testing.assert_array_equal(ds[0][:, 0], ds[0][:, 1])


@pytest.mark.parametrize("neq", (1, 2, 3))
def test_ref1d_equilibria_are_exact_fixed_points(neq):
    p = po.make_params_1d(96, nequilibrium=neq)
    x = (np.arange(96) + 0.5) / 96
    w = np.zeros((3, 96), order="F")
    lib = po.lib()
    lib.orc_ref1d_equilibrium(p.gamma, neq, 96, po._p(np.ascontiguousarray(x)), po._p(w))
    if neq == 3:   # polytropic: p = rho^gamma exactly in the closed form
        np.testing.assert_allclose(w[2], w[0] ** p.gamma, rtol=1e-14)
    else:
        np.testing.assert_allclose(w[0], np.exp(-x), rtol=4e-16)   # numpy's exp: within 1 ulp of libm
        np.testing.assert_array_equal(w[0], w[2])
    assert np.count_nonzero(po.ref1d_update(p, po.conservative_1d(p.gamma, w), w)) == 0


# ---- the configured scheme -------------------------------------------------------------
SCHEMES_1D = [(r, bc) for r in (po.RECON_FIRST_ORDER, po.RECON_MUSCL_MINMOD)
              for bc in (po.BC1D_ZERO_GRADIENT, po.BC1D_WALL)]


@pytest.mark.parametrize("scheme", SCHEMES_1D)
@pytest.mark.parametrize("neq,g", [(2, 1.0), (3, 1.0), (2, 2.5), (3, 0.7)])
def test_fv1d_rest_is_exact(scheme, neq, g):
    r, bc = scheme
    p = po.make_params_1d(77, recon=r, bc=bc, nequilibrium=neq, g=g, gamma=1.4, integrator=po.SSPRK3)
    u, w = po.perturbed_state_1d(p, amp_bump=0.0)
    assert np.count_nonzero(po.fv1d_residual(p, u, w)) == 0
    u2, t, n, _ = po.fv1d_advance(p, u, w, t_end=10.0, max_steps=60)
    assert n == 60 and np.array_equal(u2, u)


def test_fv1d_face_flux_is_consistent():
    p = po.make_params_1d(16, gamma=1.4)
    rng = np.random.default_rng(5)
    for _ in range(50):
        ueq = np.array([rng.uniform(0.5, 2), 0.0, rng.uniform(1, 4)])
        d = np.array([rng.normal() * 0.1, rng.normal() * 0.2, rng.normal() * 0.1])
        f = po.fv1d_face_flux(p, ueq, d, d)
        np.testing.assert_array_equal(f, po.ref1d_flux(p.gamma, ueq + d) - po.ref1d_flux(p.gamma, ueq))


def test_fv1d_walls_mirror_and_conserve_mass():
    """Without gravity a pulse centred in the box stays mirror symmetric
    (rho(x) = rho(1-x), m(x) = -m(1-x)), and walls keep the mass constant."""
    p = po.make_params_1d(120, recon=po.RECON_MUSCL_MINMOD, bc=po.BC1D_WALL, g=0.0, gamma=1.4,
                          integrator=po.SSPRK3)
    u, w = po.perturbed_state_1d(p, amp_bump=0.3, x0=0.5, width=0.05)
    u2, t, n, _ = po.fv1d_advance(p, u, w, t_end=0.4)
    assert t == 0.4
    np.testing.assert_allclose(u2[0], u2[0][::-1], rtol=0, atol=1e-13)
    np.testing.assert_allclose(u2[1], -u2[1][::-1], rtol=0, atol=1e-13)
    assert abs(u2[0].sum() - u[0].sum()) <= 1e-13 * u[0].sum()
    # with gravity the walls still conserve mass
    p.g, p.nequilibrium = 1.0, 2
    u, w = po.perturbed_state_1d(p, amp_bump=1e-2, x0=0.3)
    u2, _, _, _ = po.fv1d_advance(p, u, w, t_end=0.5)
    assert abs(u2[0].sum() - u[0].sum()) <= 1e-13 * u[0].sum()


def _pulse_error(nx, recon, integ=po.SSPRK3, cfl=0.4):
    p = po.make_params_1d(nx, recon=recon, bc=po.BC1D_WALL, g=1.0, gamma=1.4, integrator=integ, cfl=cfl)
    u, w = po.perturbed_state_1d(p, amp_bump=1e-3, x0=0.5, width=0.08)
    u2, _, _, _ = po.fv1d_advance(p, u, w, t_end=0.15)
    return u2


def _coarsen(a, f):
    return a.reshape(a.shape[0], -1, f).mean(axis=2)


@pytest.mark.parametrize("recon,lo,hi", [(po.RECON_FIRST_ORDER, 0.7, 1.3), (po.RECON_MUSCL_MINMOD, 1.6, 2.6)])
def test_fv1d_spatial_order(recon, lo, hi):
    """Smooth small pulse: the cell-average error against a 16x finer run
    shrinks at the scheme's order."""
    ref = _pulse_error(1280, recon)
    errs = []
    for n in (80, 160, 320):
        u = _pulse_error(n, recon)
        errs.append(np.mean(np.abs(u[2] - _coarsen(ref, 1280 // n)[2])))
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all((orders > lo) & (orders < hi)), (errs, orders)


@pytest.mark.parametrize("integ,order", [(po.RK2_HEUN, 2), (po.SSPRK3, 3)])
def test_fv1d_time_order(integ, order):
    """Same grid, CFL number halved: the change shrinks as cfl^order."""
    a, b, c = (_pulse_error(200, po.RECON_MUSCL_MINMOD, integ, cfl) for cfl in (0.4, 0.2, 0.1))
    r = np.log2(np.max(np.abs(a - b)) / np.max(np.abs(b - c)))
    assert abs(r - order) < 0.45, r


def test_fv1d_sod_converges_to_exact_solution():
    errs = []
    for nx in (200, 400):
        p = po.make_params_1d(nx, recon=po.RECON_MUSCL_MINMOD, bc=po.BC1D_WALL, g=0.0, gamma=1.4,
                              integrator=po.SSPRK3)
        x = (np.arange(nx) + 0.5) / nx
        w = po.equilibrium_1d(p, x)
        ws = np.zeros((3, nx), order="F")
        ws[0] = np.where(x < 0.5, 1.0, 0.125)
        ws[2] = np.where(x < 0.5, 1.0, 0.1)
        u, t, n, _ = po.fv1d_advance(p, po.conservative_1d(p.gamma, ws), w, t_end=0.2)
        assert t == 0.2
        exact, _, _ = exact_riemann_density(1.0, 0.0, 1.0, 0.125, 0.0, 0.1, 1.4, (x - 0.5) / t)
        errs.append(np.mean(np.abs(u[0] - exact)))
    assert errs[0] < 0.02 and errs[1] < 0.7 * errs[0], errs


@pytest.mark.parametrize("recon", (po.RECON_FIRST_ORDER, po.RECON_MUSCL_MINMOD))
def test_fv1d_equals_2d_oracle_on_y_uniform_problem(recon):
    """Without gravity, a 2D problem uniform in y with reflecting walls is the
    1D problem in every row: the 1D oracle (Rusanov, SSP-RK2) and the 2D
    oracle agree to rounding over 25 steps of the same dt."""
    nx = 90
    p1 = po.make_params_1d(nx, recon=recon, bc=po.BC1D_WALL, g=0.0, gamma=1.4, cfl=0.3)
    p2 = po.make_params(nx, 4, recon=recon, flux=po.FLUX_RUSANOV, bc=po.BC_REFLECT, integrator=po.RK2_HEUN,
                        gamma=1.4, rho0=1.0, p0=1.0, g=0.0, cfl=0.6)
    u1, w1 = po.perturbed_state_1d(p1, amp_bump=0.2, x0=0.35, width=0.06, amp_noise=1e-3)
    u2 = np.zeros((4, nx, 4), order="F")
    w2 = np.zeros((4, nx, 4), order="F")
    for j in range(4):
        u2[[0, 1, 3], :, j] = u1
        w2[[0, 1, 3], :, j] = w1
    a, _, n1, dt1 = po.fv1d_advance(p1, u1, w1, max_steps=25)
    b, _, n2, dt2 = po.fv2d_advance(p2, u2, w2, max_steps=25)
    assert n1 == n2 == 25 and abs(dt1 - dt2) <= 1e-15 * dt1
    for j in range(4):
        np.testing.assert_allclose(b[[0, 1, 3], :, j], a, rtol=0, atol=1e-13)
    assert np.max(np.abs(b[2])) < 1e-15
