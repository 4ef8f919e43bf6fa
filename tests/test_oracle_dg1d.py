"""The CPU oracle of the reference's 1D modal DG (dg_with_source.f90 with
legendre.f90): bit for bit against the fixtures written from the flang build
of the reference, live against oracle/_ref when present, and against what the
mathematics fixes (the rest state, the Gauss rules, the order of accuracy and
the scaling of the residual)."""
import os

import numpy as np
import pytest

import pyoracle as po
import ref_runner as rr

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EPS = np.finfo(float).eps


def load(name):
    with np.load(os.path.join(GOLD, name + ".npz")) as z:
        return {k: z[k] for k in z.files}


def test_legendre_matches_reference_golden_bitwise():
    """legendre / legendre_prime (legendre.f90:1-48) for n = 0..6, including
    the FP32 square roots of the normalisation and the clamp at +-1.5."""
    g = load("dg1d_legendre")
    for k in range(7):
        for i, x in enumerate(g["x"]):
            assert po.dg1d_legendre(x, k) == g["P"][k, i], (k, x)
            assert po.dg1d_legendre_prime(x, k) == g["dP"][k, i], (k, x)


def test_gl_quadrature_matches_reference_golden_bitwise():
    """gl_quadrature (legendre.f90:77-127): FP32 literals for n <= 3, Newton
    roots with the weight 2(2n+1)/((1-x^2)P_n'^2) as written for n >= 4."""
    g = load("dg1d_gl_quadrature")
    for n in range(1, 7):
        x, w = po.dg1d_gl_quadrature(n)
        np.testing.assert_array_equal(x, g[f"x{n}"])
        np.testing.assert_array_equal(w, g[f"w{n}"])


def test_gl_quadrature_points_are_the_gauss_points():
    """The points are the Gauss-Legendre nodes (to FP32 rounding for n <= 3,
    to FP64 for the Newton roots); the weights are Gauss's for n <= 3 and
    (2n+1) times Gauss's for n >= 4, which is how the reference writes them."""
    for n in range(1, 7):
        x, w = po.dg1d_gl_quadrature(n)
        xr, wr = np.polynomial.legendre.leggauss(n)
        tol = 1e-7 if n <= 3 else 1e-14
        np.testing.assert_allclose(np.sort(x), xr, atol=tol)
        scale = 1.0 if n <= 3 else 2 * n + 1
        np.testing.assert_allclose(w[np.argsort(x)], scale * wr, rtol=1e-7 if n <= 3 else 1e-13)


def test_basis_is_orthonormal_up_to_degree_2():
    """phi_0..phi_2 are orthonormal on [-1,1] (to the FP32 square roots);
    phi_3 is the plain P_3, of norm 2/7."""
    xg, wg = np.polynomial.legendre.leggauss(8)
    P = np.array([[po.dg1d_legendre(x, i) for x in xg] for i in range(4)])
    G = (P * wg) @ P.T
    np.testing.assert_allclose(G[:3, :3], np.eye(3), atol=2e-7)
    assert abs(G[3, 3] - 2 / 7) < 1e-14


def test_update_exact_delta_matches_reference_golden_bitwise():
    g = load("dg1d_update_exact_delta_128")
    p = po.make_dg1d_params(128)
    np.testing.assert_array_equal(po.dg1d_equilibrium_nodes(p), g["u_eq"])
    for k in ("init", "rand", "strong", "zero"):
        np.testing.assert_array_equal(po.dg1d_update_exact_delta(p, g["du_" + k], g["u_eq"]), g["dudt_" + k],
                                      err_msg=k)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 6])
@pytest.mark.parametrize("riemann", [po.DG1D_RIEMANN_LLF, po.DG1D_RIEMANN_HLLC])
@pytest.mark.parametrize("source", [1, 2])
def test_rest_state(n, riemann, source):
    """Zero modes: the volume and source differences vanish exactly; the faces
    leave Rusanov's flux of two equal states, F(u) exactly, and HLLC's star
    pressure (DR pL + DL pR)/(DL + DR), which rounds (the reference's own rest
    residual is 2.8e-14, not 0: see the fixture)."""
    p = po.make_dg1d_params(64, n=n, riemann=riemann, source=source)
    d = po.dg1d_update_exact_delta(p, np.zeros((3, n, 64)), po.dg1d_equilibrium_nodes(p))
    if riemann == po.DG1D_RIEMANN_LLF:
        assert np.count_nonzero(d) == 0
    else:
        assert np.max(np.abs(d)) <= 16 * EPS * 64 * 10


@pytest.mark.parametrize("riemann", [po.DG1D_RIEMANN_LLF, po.DG1D_RIEMANN_HLLC])
def test_residual_is_half_the_dg_rate_at_third_order(riemann):
    """For the orthonormal basis the textbook DG rate of mode 0 is
    int phi_0 (-(dF/dx) + S) dxi; the reference's is half of it (its factors
    1/dx and 1/2 where 2/dx and 1 would be), so d(mode 0)/dt / phi_0 tends to
    the cell average of -(F(U_eq+d) - F(U_eq))_x + S(U_eq+d) - S(U_eq), at
    third order for n = 3.  A sign error, a lost term or the textbook factor
    leaves an O(1) error."""
    def delta(x):
        return np.stack([0.05 * np.sin(2 * np.pi * x), 0.04 * np.cos(2 * np.pi * x),
                         0.06 * np.sin(4 * np.pi * x) + 0.02])

    def flux(U, g):
        r, m, E = U
        v = m / r
        pr = (g - 1) * (E - 0.5 * r * v * v)
        return np.stack([m, m * v + pr, (E + pr) * v])

    def src(U):
        return np.stack([0 * U[0], -U[0], -U[1]])

    errs = []
    for nx in (16, 32, 64):
        p = po.make_dg1d_params(nx, riemann=riemann)
        g, dx = p.gamma, 1.0 / nx

        def ueq(y):
            e = np.exp(-y)
            return np.stack([e, 0 * e, e / (g - 1)])

        xg, wg = np.polynomial.legendre.leggauss(6)
        phg = np.array([[po.dg1d_legendre(v, i) for v in xg] for i in range(3)])
        X = ((np.arange(nx) + 0.5) * dx)[None, :] + dx / 2 * xg[:, None]
        c = np.einsum("vjc,ij,j->vic", delta(X), phg, wg)           # int delta phi_i dxi
        d = po.dg1d_update_exact_delta(p, c, po.dg1d_equilibrium_nodes(p))
        h = 1e-6

        def dF(y):
            return flux(ueq(y) + delta(y), g) - flux(ueq(y), g)

        R = -(dF(X + h) - dF(X - h)) / (2 * h) + src(ueq(X) + delta(X)) - src(ueq(X))
        avg = 0.5 * np.einsum("vjc,j->vc", R, wg)
        phi0 = po.dg1d_legendre(0.0, 0)
        errs.append(np.max(np.abs(d[:, 0, :] / phi0 - avg)[:, 2:-2]))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert errs[0] < 1e-2 and np.all(rates > 2.5), (errs, rates)


def test_end_cells_are_frozen_and_bc_free():
    """dudt of cells 1 and nx is zero (:2027-2028), and no other cell sees the
    two boundary faces."""
    p = po.make_dg1d_params(32)
    rng = np.random.default_rng(3)
    du = rng.normal(size=(3, 3, 32)) * 1e-2
    d = po.dg1d_update_exact_delta(p, du, po.dg1d_equilibrium_nodes(p))
    assert np.count_nonzero(d[:, :, [0, -1]]) == 0
    du2 = du.copy()
    du2[:, :, 0] += 0.05
    d2 = po.dg1d_update_exact_delta(p, du2, po.dg1d_equilibrium_nodes(p))
    np.testing.assert_array_equal(d2[:, :, 2:], d[:, :, 2:])


@pytest.mark.skipif(not rr.available("dg1d"), reason="oracle/_ref not built")
@pytest.mark.parametrize("seed", [1, 2])
def test_update_exact_delta_live_against_reference(seed):
    g = load("dg1d_update_exact_delta_128")
    du = np.asfortranarray(np.random.default_rng(seed).normal(size=(3, 3, 128)) * 0.05)
    ref = rr.run("dg1d_update_exact_delta", delta_u=du, u_eq=g["u_eq"])["dudt"]
    np.testing.assert_array_equal(po.dg1d_update_exact_delta(po.make_dg1d_params(128), du, g["u_eq"]), ref)


def final_error(p, u_init):
    """The reference program's closing 'error' (:391-406): max over cells of
    |p - p_eq| at the first Gauss point of the last reconstruction."""
    chsi, _ = po.dg1d_gl_quadrature(p.n)
    dx = p.boxlen / p.nx
    x = (np.arange(1, p.nx + 1) - 0.5) * dx + dx / 2.0 * chsi[0]
    U = u_init[:, 0, :]
    v = U[1] / U[0]
    pr = (p.gamma - 1.0) * (U[2] - 0.5 * U[0] * (v * v))
    return np.max(np.abs(pr - np.exp(-x))), x, pr - np.exp(-x)


def test_run_rki_matches_reference_program_bitwise():
    """The reference's own 1D DG run (dg_commons.f90: n = 3, nx = 128, HLLC,
    ninit = 8 with pert = 1e-8, integrator 'RKi', tend = 0.2): the oracle's
    restatement of the main program takes the same 236 steps with every t and
    dt equal bit for bit, ends at the same 'error' bit for bit, and writes the
    same recon99999.dat to its six printed digits."""
    g = load("dg1d_run_rki_128")
    p = po.make_dg1d_params(128)
    d, e, u = po.dg1d_initial(p)
    d, u, t, k, ts, dts = po.dg1d_run_rki(p, d, e, u, 0.2)
    assert k == len(g["t"]) == 236
    np.testing.assert_array_equal(ts, g["t"])
    np.testing.assert_array_equal(dts, g["dt"])
    err, x, dp = final_error(p, u)
    assert err == float(g["error"])
    np.testing.assert_allclose(x, g["recon"][:, 0], rtol=6e-6)
    # p - p_eq at six digits; numpy's exp in final_error may differ from libm's by
    # an ulp of the O(1) pressure, hence the absolute part
    np.testing.assert_allclose(dp, g["recon"][:, 1], rtol=6e-6, atol=4 * EPS)


def test_rki_coefficients_are_the_ssp54_scheme():
    """The fifteen literals (rounded to FP32) satisfy the order conditions of a
    fourth-order method to FP32 accuracy when written as a Butcher tableau."""
    c = po.dg1d_rki_coeffs()
    A = np.zeros((5, 5))
    b = np.zeros(5)
    # Shu-Osher form u^(i) = sum_k alpha_ik u^(k) + beta_ik dt L(u^(k)), rows of
    # (alpha, beta) in terms of stage vectors u0..u4
    alpha = np.zeros((6, 5)); beta = np.zeros((6, 5))
    alpha[1, 0], beta[1, 0] = 1.0, c[0]
    alpha[2, 0], alpha[2, 1], beta[2, 1] = c[1], c[2], c[3]
    alpha[3, 0], alpha[3, 2], beta[3, 2] = c[4], c[5], c[6]
    alpha[4, 0], alpha[4, 3], beta[4, 3] = c[7], c[8], c[9]
    alpha[5, 2], alpha[5, 3], beta[5, 3], alpha[5, 4], beta[5, 4] = c[10], c[11], c[12], c[13], c[14]
    # Butcher coefficients: K[i] = sum_k alpha_ik K[k] + beta_ik e_k
    K = np.zeros((6, 5))
    for i in range(1, 6):
        K[i] = sum(alpha[i, k] * K[k] for k in range(i)) + beta[i]
    A, b = K[:5], K[5]
    cc = A.sum(axis=1)
    for lhs, rhs in [(b.sum(), 1), (b @ cc, 1 / 2), (b @ cc ** 2, 1 / 3), (b @ A @ cc, 1 / 6),
                     (b @ cc ** 3, 1 / 4), (b @ (cc * (A @ cc)), 1 / 8), (b @ A @ cc ** 2, 1 / 12),
                     (b @ A @ A @ cc, 1 / 24)]:
        assert abs(lhs - rhs) < 1e-6, (lhs, rhs)
