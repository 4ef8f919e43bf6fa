"""The 2D modal DG oracle.  The reference's building blocks (Legendre basis,
Gauss-Legendre quadrature, tensor projections of 2d/) against the fixtures
written from the flang-built reference, bit for bit; the configured P_k
deviation-form scheme against what the mathematics fixes: closed-form nodes
and weights, the mass identity, exact rest, degree 0 equal to the first-order
FV oracle, convergence at order k+1, wall mass conservation, x<->y symmetry,
and a TVB limiter that keeps means and leaves smooth data alone."""
import os

import numpy as np
import pytest

import pyoracle as po
import ref_runner as rr

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(GOLD, name + ".npz")) as z:
        return {k: z[k] for k in z.files}


# ---- the reference's building blocks -------------------------------------------------
def test_legendre_matches_reference_golden_bitwise():
    g = load("dg2d_legendre")
    P = np.array([[po.dg_legendre(x, n) for n in range(7)] for x in g["x"]])
    dP = np.array([[po.dg_legendre_prime(x, n) for n in range(7)] for x in g["x"]])
    np.testing.assert_array_equal(P, g["P"])
    np.testing.assert_array_equal(dP, g["dP"])
    # 2d/legendre.f90:5 clamps the argument: P_n(1.5) == P_n(1) = sqrt(2n+1)
    np.testing.assert_allclose(P[-1], np.sqrt(2 * np.arange(7) + 1.0), rtol=1e-15)


def test_gl_quadrature_matches_reference_golden_and_closed_forms():
    g = load("dg2d_gl_quadrature")
    for n in range(1, 7):
        x, w = po.dg_gl_quadrature(n)
        np.testing.assert_array_equal(x, g[f"x{n}"])
        np.testing.assert_array_equal(w, g[f"w{n}"])
        xr, wr = np.polynomial.legendre.leggauss(n)
        np.testing.assert_allclose(np.sort(x), xr, atol=2e-15)
        np.testing.assert_allclose(w[np.argsort(x)], wr, atol=2e-15)
    x, w = po.dg_gl_quadrature(3)
    np.testing.assert_allclose(sorted(x), [-np.sqrt(0.6), 0.0, np.sqrt(0.6)], atol=1e-15)
    np.testing.assert_allclose(sorted(w), [5 / 9, 5 / 9, 8 / 9], atol=1e-15)


def test_tensor_projections_match_reference_golden_bitwise():
    g = load("dg2d_projections")
    np.testing.assert_array_equal(po.ref_dg2d_modes_from_nodes(g["a"]), g["modes_from_nodes"])
    np.testing.assert_array_equal(po.ref_dg2d_nodes_from_modes(g["a"]), g["nodes_from_modes"])
    # with as many nodes as modes the two are inverse to each other
    back = po.ref_dg2d_modes_from_nodes(po.ref_dg2d_nodes_from_modes(g["a"]))
    np.testing.assert_allclose(back, g["a"], atol=1e-14)


@pytest.mark.skipif(not rr.available("dg2d"), reason="oracle/_ref not built")
def test_tensor_projections_match_live_reference():
    a = np.asfortranarray(np.random.default_rng(5).normal(size=(4, 8, 8, 2, 2)))
    np.testing.assert_array_equal(po.ref_dg2d_modes_from_nodes(a), rr.run("dg2d_modes_from_nodes", a=a)["r"])
    np.testing.assert_array_equal(po.ref_dg2d_nodes_from_modes(a), rr.run("dg2d_nodes_from_modes", a=a)["r"])


# ---- the configured scheme -------------------------------------------------------------
def rand_dev(p, scale=1e-3, seed=0):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(scale * rng.normal(size=(4, p.nx, p.ny, po.dg_nmodes(p.degree))))


@pytest.mark.parametrize("degree", (0, 1, 2))
def test_projection_of_nodal_values_is_identity(degree):
    """(1/4) sum_q w_q phi_m phi_n = delta_mn on (k+1)^2 Gauss points."""
    p = po.make_dg_params(5, 4, degree=degree)
    d = rand_dev(p, 1.0)
    np.testing.assert_allclose(po.dg2d_project(p, po.dg2d_nodal(p, d)), d, atol=2e-15)


@pytest.mark.parametrize("degree", (0, 1, 2))
@pytest.mark.parametrize("flux", (po.FLUX_RUSANOV, po.FLUX_HLL, po.FLUX_HLLC))
def test_rest_is_exact(degree, flux):
    p = po.make_dg_params(7, 6, degree=degree, flux=flux, g=1.7, rho0=1.2, p0=0.9)
    d = np.zeros((4, 7, 6, po.dg_nmodes(degree)), order="F")
    assert np.count_nonzero(po.dg2d_residual(p, d)) == 0
    d2, t, n, _ = po.dg2d_advance(p, d, max_steps=5)
    assert n == 5 and np.count_nonzero(d2) == 0


@pytest.mark.parametrize("flux", (po.FLUX_RUSANOV, po.FLUX_HLL, po.FLUX_HLLC))
def test_degree_zero_is_first_order_fv(flux):
    """P0 with one Gauss point per direction is the first-order FV scheme with
    walls: the same face fluxes at the face midpoints and the same source."""
    nx, ny = 9, 7
    pd = po.make_dg_params(nx, ny, degree=0, flux=flux, g=1.3, rho0=1.1, p0=0.8)
    pf = po.make_params(nx, ny, recon=po.RECON_FIRST_ORDER, flux=flux, bc=po.BC_REFLECT, gamma=1.4, rho0=1.1,
                        p0=0.8, g=1.3)
    d = rand_dev(pd, 1e-2, seed=3)
    w = po.equilibrium_primitive(pf)
    u = po.conservative(pf.gamma, w) + d[:, :, :, 0]
    np.testing.assert_allclose(po.dg2d_residual(pd, d)[:, :, :, 0], po.fv2d_residual(pf, u, w), rtol=0, atol=1e-13)


def test_transpose_symmetry():
    p = po.make_dg_params(6, 6, degree=2, g=1.0)
    d = rand_dev(p, 1e-2, seed=4)
    # transposing the grid swaps x<->y, the two momenta and modes (1,0)<->(0,1), (2,0)<->(0,2)
    perm = [0, 2, 1, 5, 4, 3]
    dt_ = d.transpose(0, 2, 1, 3)[[0, 2, 1, 3]][:, :, :, perm]
    L = po.dg2d_residual(p, d)
    Lt = po.dg2d_residual(p, np.asfortranarray(dt_))
    np.testing.assert_allclose(Lt, L.transpose(0, 2, 1, 3)[[0, 2, 1, 3]][:, :, :, perm], rtol=0, atol=1e-13)


def test_walls_conserve_mass():
    p = po.make_dg_params(12, 10, degree=2, g=1.0, limiter=po.DG_LIMITER_NONE)
    d = po.dg2d_fourier_state(p, amp=1e-2)
    d2, _, n, _ = po.dg2d_advance(p, d, max_steps=8)
    assert n == 8
    assert abs(d2[0, :, :, 0].sum() - d[0, :, :, 0].sum()) < 1e-15


def _acoustic(n, degree):
    """Pressure wave dp = 1e-3 cos(pi x) cos(pi y) on the g = 1 atmosphere
    (compatible with the walls: its even extension is smooth), t = 0.05."""
    p = po.make_dg_params(n, n, degree=degree, g=1.0, limiter=po.DG_LIMITER_NONE, cfl=0.5)
    modes = (np.array([1.0, 1.0]), np.array([0.5, 0.5]), np.array([0.5, -0.5]), np.array([0.0, 0.0]))
    d = po.dg2d_fourier_state(p, amp=1e-3, modes=modes)
    d2, t, _, _ = po.dg2d_advance(p, d, t_end=0.05)
    assert abs(t - 0.05) < 1e-15
    return d2[:, :, :, 0]   # cell means


_REF = {}


@pytest.mark.parametrize("degree,lo", [(1, 2.2), (2, 2.9)])
def test_convergence_order(degree, lo):
    """Cell means against a P2 run 4-8x finer: the energy error falls at
    order about k+1 (measured 2.7 for P1, 3.4 for P2)."""
    if "ref" not in _REF:
        _REF["ref"] = _acoustic(64, 2)
    ref = _REF["ref"]
    errs = []
    for n in (8, 16):
        f = 64 // n
        coarse = ref.reshape(4, n, f, n, f).mean(axis=(2, 4))
        errs.append(np.mean(np.abs(_acoustic(n, degree)[3] - coarse[3])))
    order = np.log2(errs[0] / errs[1])
    assert order > lo, (errs, order)


def test_tvb_limiter_keeps_means_and_smooth_data():
    p = po.make_dg_params(16, 16, degree=2, tvb_m=50.0)
    d = po.dg2d_fourier_state(p, amp=1e-4)
    np.testing.assert_array_equal(po.dg2d_limit(p, d), d)       # slopes below M dx^2: untouched
    # a pressure step: limited cells keep their means, lose their quadratic modes
    X, Y = po.dg2d_points(p)
    nodal = np.zeros((4, 16, 16, 9), order="F")
    nodal[3] = np.where(X < 0.5, 1.0, 0.0)
    nodal[0] = np.where(X + Y < 1.0, 0.5, 0.0)
    d = po.dg2d_project(p, nodal)
    lim = po.dg2d_limit(p, d)
    np.testing.assert_array_equal(lim[:, :, :, 0], d[:, :, :, 0])
    changed = np.any(lim != d, axis=3)
    assert changed.any()
    assert np.all(lim[:, :, :, 3:][changed] == 0.0)
    p0 = po.make_dg_params(16, 16, degree=2, tvb_m=0.0)
    lim0 = po.dg2d_limit(p0, d)
    # TVD (M = 0): linear modes never exceed the neighbour differences
    s3 = np.sqrt(3.0)
    m = lim0[3, :, :, 0]
    dxp = np.diff(m, axis=0)
    a = s3 * lim0[3, 1:-1, :, 1]
    assert np.all(np.abs(a) <= np.maximum(np.abs(dxp[1:]), np.abs(dxp[:-1])) + 1e-15)


# ---- the positivity limiter (2d/limiters.f90:312-654) ---------------------------------
def test_reference_positivity_matches_golden_bitwise():
    """compute_positivity as the reference compiles it (its shipped limiter
    'ONP'; 8x8 cells of tensor modes, mx = my = 2): the oracle's restatement is
    bit for bit the reference's output, on a state where most cells are limited."""
    g = load("dg2d_positivity")
    out = po.ref_dg2d_positivity(g["u"], gamma=float(g["gamma"]), eps=float(g["eps"]))
    np.testing.assert_array_equal(out, g["u_limited"])
    assert np.any(out != g["u"])
    np.testing.assert_array_equal(out[:, :, :, 0, 0], g["u"][:, :, :, 0, 0])   # means untouched


def test_solve_for_t_puts_the_pressure_at_eps():
    """The root of solve_for_t moves a point of negative pressure towards the
    cell mean until its pressure is eps (the quadratic of (1-t) u_avg + t u)."""
    g = 1.4
    eps = float(np.float32(1e-10))
    rng = np.random.default_rng(312)
    for _ in range(200):
        ua = np.array([rng.uniform(0.5, 2), rng.uniform(-1, 1), rng.uniform(-1, 1), 0.0])
        ua[3] = rng.uniform(0.3, 2) / (g - 1) + 0.5 * (ua[1] ** 2 + ua[2] ** 2) / ua[0]
        u = np.array([rng.uniform(0.2, 2), rng.uniform(-3, 3), rng.uniform(-3, 3), 0.0])
        u[3] = 0.5 * (u[1] ** 2 + u[2] ** 2) / u[0] - rng.uniform(0.01, 1) / (g - 1)   # p < 0
        t = po.ref_dg2d_solve_for_t(g, eps, u, ua)
        assert 0.0 <= t <= 1.0
        w = (1 - t) * ua + t * u
        pr = (g - 1) * (w[3] - 0.5 * (w[1] ** 2 + w[2] ** 2) / w[0])
        assert abs(pr - eps) < 1e-9 * (1 + abs(ua[3]) + abs(u[3]))


def set_values(p, F):
    """Full-state modes F (4,nx,ny,6) at the limiter's set: (x_gll, y_gl),
    (x_gl, y_gll) and (x_gl, y_gl), Gauss-Lobatto -1, 0, 1 and the three Gauss
    points."""
    xq, _ = po.dg_gl_quadrature(3)
    xg = np.array([-1.0, 0.0, 1.0])
    mi, mj = [0, 1, 0, 2, 1, 0], [0, 0, 1, 0, 1, 2]
    ph = lambda n, x: po.dg_legendre(x, n)
    out = []
    for q in range(3):
        for c in range(3):
            for (x, y) in ((xg[c], xq[q]), (xq[q], xg[c]), (xq[q], xq[c])):
                out.append(sum(F[..., m] * ph(mi[m], x) * ph(mj[m], y) for m in range(6)))
    return np.stack(out, axis=-1)


def equilibrium_modes(p):
    X, Y = po.dg2d_points(p)
    k = p.rho0 * p.g / p.p0
    e = np.exp(-k * (X + Y))
    nodal = np.zeros((4,) + e.shape, order="F")
    nodal[0] = p.rho0 * e
    nodal[3] = p.p0 * e / (p.gamma - 1)
    return po.dg2d_project(p, nodal)


def test_positivity_limiter_keeps_means_and_makes_the_set_positive():
    """The configured limiter on a state with vacuum-like cells: the means of
    every variable are untouched bit for bit, and afterwards density and
    pressure of the full state are at least about the cell's thresholds (1e-5
    of its equilibrium density and pressure) at every set point."""
    p = po.make_dg_params(12, 12, degree=2, limiter=po.DG_LIMITER_POSITIVITY)
    rng = np.random.default_rng(478)
    E = equilibrium_modes(p)
    d = np.asfortranarray(rng.normal(size=(4, 12, 12, 6)) * 0.6 * np.abs(E[..., :1]))
    d[..., 0] = 0.0                               # means: the equilibrium's, positive
    before = set_values(p, E + d)
    assert (before[0] < 0).any()
    lim = po.dg2d_limit(p, d)
    np.testing.assert_array_equal(lim[..., 0], d[..., 0])
    after = set_values(p, E + lim)
    r = after[0]
    pr = (p.gamma - 1) * (after[3] - 0.5 * (after[1] ** 2 + after[2] ** 2) / r)
    eq = set_values(p, E)
    assert np.all(r >= 1e-5 * E[0, :, :, 0][..., None] * (1 - 1e-9))
    assert np.all(pr >= 1e-5 * (p.gamma - 1) * E[3, :, :, 0][..., None] * (1 - 1e-6))
    assert eq[0].min() > 0
    assert np.any(lim != d)


def test_positivity_limiter_leaves_positive_states_alone():
    p = po.make_dg_params(16, 16, degree=2, limiter=po.DG_LIMITER_POSITIVITY)
    d = po.dg2d_fourier_state(p, amp=1e-2)
    np.testing.assert_array_equal(po.dg2d_limit(p, d), d)
    z = np.zeros((4, 16, 16, 6), order="F")
    np.testing.assert_array_equal(po.dg2d_limit(p, z), z)


def test_tvb_then_positivity():
    p3 = po.make_dg_params(12, 12, degree=2, limiter=po.DG_LIMITER_TVB_POSITIVITY, tvb_m=0.0)
    p1 = po.make_dg_params(12, 12, degree=2, limiter=po.DG_LIMITER_TVB, tvb_m=0.0)
    p2 = po.make_dg_params(12, 12, degree=2, limiter=po.DG_LIMITER_POSITIVITY, tvb_m=0.0)
    E = equilibrium_modes(p3)
    d = np.asfortranarray(np.random.default_rng(5).normal(size=(4, 12, 12, 6)) * 0.5 * np.abs(E[..., :1]))
    np.testing.assert_array_equal(po.dg2d_limit(p3, d), po.dg2d_limit(p2, po.dg2d_limit(p1, d)))


# ---- characteristic limiting -----------------------------------------------------------
def euler_flux(gamma, U, direction):
    r, mx, my, E = U
    vx, vy = mx / r, my / r
    p = (gamma - 1) * (E - 0.5 * r * (vx * vx + vy * vy))
    if direction == 0:
        return np.array([mx, mx * vx + p, my * vx, (E + p) * vx])
    return np.array([my, mx * vy, my * vy + p, (E + p) * vy])


def test_eigenvectors_diagonalise_the_flux_jacobian():
    """R's columns are eigenvectors of dF/dU (by central differences) for the
    speeds vn-c, vn, vn, vn+c, and L is R's inverse, in both directions."""
    g = 1.4
    rng = np.random.default_rng(21)
    for _ in range(20):
        rho, vx, vy, pr = rng.uniform(0.5, 2), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.3, 2)
        U = np.array([rho, rho * vx, rho * vy, pr / (g - 1) + 0.5 * rho * (vx * vx + vy * vy)])
        c = np.sqrt(g * pr / rho)
        for d in (0, 1):
            R, L = po.euler_eigenvectors(g, d, U)
            np.testing.assert_allclose(L @ R, np.eye(4), atol=1e-13)
            h = 1e-6
            A = np.stack([(euler_flux(g, U + h * e, d) - euler_flux(g, U - h * e, d)) / (2 * h)
                          for e in np.eye(4)], axis=1)
            vn = vx if d == 0 else vy
            np.testing.assert_allclose(A @ R, R * np.array([vn - c, vn, vn, vn + c]), atol=1e-7)


def test_characteristic_limiter_basics():
    """Smooth data below M dx^2 and a resting atmosphere are left alone bit
    for bit; on a step every mean stays, a limited cell keeps only its
    (characteristically limited) linear modes, and at M = 0 the limited x
    slope in characteristic variables is within the neighbour differences."""
    p = po.make_dg_params(16, 16, degree=2, tvb_m=50.0, limiter=po.DG_LIMITER_CHARACTERISTIC)
    d = po.dg2d_fourier_state(p, amp=1e-4)
    np.testing.assert_array_equal(po.dg2d_limit(p, d), d)
    z = np.zeros((4, 16, 16, 6), order="F")
    np.testing.assert_array_equal(po.dg2d_limit(p, z), z)
    X, Y = po.dg2d_points(p)
    nodal = np.zeros((4, 16, 16, 9), order="F")
    nodal[3] = np.where(X < 0.5, 0.3, 0.0)
    nodal[0] = np.where(X + Y < 1.0, 0.2, 0.0)
    nodal[1] = 0.1 * np.sin(3 * Y)
    d = po.dg2d_project(p, nodal)
    p0 = po.make_dg_params(16, 16, degree=2, tvb_m=0.0, limiter=po.DG_LIMITER_CHARACTERISTIC)
    lim = po.dg2d_limit(p0, d)
    np.testing.assert_array_equal(lim[..., 0], d[..., 0])
    changed = np.any(lim != d, axis=(0, 3))
    assert changed.any()
    assert np.all(lim[:, changed][:, :, 3:] == 0.0)
    E = equilibrium_modes(p0)
    s3 = np.sqrt(3.0)
    for i in range(1, 15):
        for j in range(16):
            ubar = E[:, i, j, 0] + d[:, i, j, 0]
            _, L = po.euler_eigenvectors(p0.gamma, 0, ubar)
            a = L @ (s3 * lim[:, i, j, 1])
            dp = L @ (d[:, i + 1, j, 0] - d[:, i, j, 0])
            dm = L @ (d[:, i, j, 0] - d[:, i - 1, j, 0])
            assert np.all(np.abs(a) <= np.maximum(np.abs(dp), np.abs(dm)) * (1 + 1e-12) + 1e-15)


def test_characteristic_limiter_on_an_entropy_wave_is_the_scalar_one():
    """A density step at rest with no pressure or velocity deviation excites
    only the entropy field, whose characteristic variable is the density
    deviation itself: the characteristic limiter then equals the scalar TVB
    limiter (to the rounding of the two transforms)."""
    pc = po.make_dg_params(16, 12, degree=2, tvb_m=0.0, limiter=po.DG_LIMITER_CHARACTERISTIC)
    ps = po.make_dg_params(16, 12, degree=2, tvb_m=0.0, limiter=po.DG_LIMITER_TVB)
    X, Y = po.dg2d_points(pc)
    nodal = np.zeros((4, 16, 12, 9), order="F")
    nodal[0] = np.where(X < 0.4, 0.2, 0.0) + np.where(Y > 0.6, 0.1, 0.0)
    d = po.dg2d_project(pc, nodal)
    lc, ls = po.dg2d_limit(pc, d), po.dg2d_limit(ps, d)
    assert np.any(ls != d)
    np.testing.assert_allclose(lc, ls, rtol=0, atol=1e-14)


# ---- the reference's SSP(5,4) (2d/benchmark_2d_dg.f90:683-710) ------------------------

def _ssp54_butcher():
    """Butcher tableau (A, b) of the Shu-Osher stages of the oracle's SSP(5,4):
    every stage is u + dt sum_j A[i, j] L(Y_j)."""
    c10, a20, a21, c21, a30, a32, c32, a40, a43, c43, a50, a52, a53, a54, c53, c54 = po.dg2d_ssp54_coeffs()
    A = np.zeros((5, 5))
    A[1, 0] = c10                                       # Y1 = w1
    A[2] = a21 * A[1]; A[2, 1] += c21                   # Y2 = w2
    A[3] = a32 * A[2]; A[3, 2] += c32                   # Y3 = w3
    A[4] = a43 * A[3]; A[4, 3] += c43                   # Y4 = w4
    b = a52 * A[2] + a53 * A[3] + a54 * A[4]
    b[3] += c53
    b[4] += c54
    return A, b


def test_ssp54_coefficients_satisfy_the_fourth_order_conditions():
    """The eight order conditions of a fourth-order Runge-Kutta method, on the
    tableau the Shu-Osher stages define; they hold to the FP32 rounding of the
    reference's default-real literals (about 1e-7), which a wrong digit or a
    dropped term anywhere breaks."""
    c = po.dg2d_ssp54_coeffs()
    assert np.array_equal(c, c.astype(np.float32).astype(np.float64))   # FP32 literals
    c10, a20, a21, c21, a30, a32, c32, a40, a43, c43, a50, a52, a53, a54, c53, c54 = c
    for s in (a20 + a21, a30 + a32, a40 + a43, a50 + a52 + a53 + a54):   # consistency
        assert abs(s - 1.0) < 1e-7
    A, b = _ssp54_butcher()
    cc = A.sum(axis=1)
    conds = [(b.sum(), 1.0), (b @ cc, 1 / 2), (b @ cc**2, 1 / 3), (b @ (A @ cc), 1 / 6),
             (b @ cc**3, 1 / 4), (b @ (cc * (A @ cc)), 1 / 8), (b @ (A @ cc**2), 1 / 12),
             (b @ (A @ (A @ cc)), 1 / 24)]
    for got, want in conds:
        assert abs(got - want) < 1e-6, conds
    # and not fifth order: the method is the five-stage fourth-order one
    assert abs(b @ cc**4 - 1 / 5) > 1e-4 or abs(b @ (A @ (A @ (A @ cc))) - 1 / 120) > 1e-4


def _ssp54_run(cfl, t_end=0.04, n=12):
    p = po.make_dg_params(n, n, degree=2, limiter=po.DG_LIMITER_NONE, integrator=po.DG_SSPRK54, cfl=cfl)
    d = po.dg2d_fourier_state(p, amp=1e-3)
    out, t, steps, _ = po.dg2d_advance(p, d, t_end=t_end)
    assert t == t_end
    return out, steps


def test_ssp54_runs_converge_at_fourth_order_in_time():
    """Same grid, dt halved twice: successive differences shrink by about 2^4
    (SSP-RK3 gives 2^3)."""
    a, na = _ssp54_run(0.8)
    b, nb = _ssp54_run(0.4)
    c, nc = _ssp54_run(0.2)
    assert na < nb < nc
    r = np.max(np.abs(a - b)) / np.max(np.abs(b - c))
    assert r > 12.0, r


def test_ssp54_keeps_rest_exact():
    p = po.make_dg_params(10, 9, degree=2, integrator=po.DG_SSPRK54, g=2.0)
    z = np.zeros((4, 10, 9, 6), order="F")
    out, t, steps, _ = po.dg2d_advance(p, z, max_steps=7)
    assert steps == 7 and np.count_nonzero(out) == 0
