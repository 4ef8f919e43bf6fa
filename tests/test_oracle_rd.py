"""The residual-distribution oracle (oracle/fvwb_oracle_rd.c, a restatement of
benchmark_rd.f90:448-547, 612-767, 346-364) pinned without the reference: the
program needs a module parameters_rd that it does not ship, so it cannot be
built.  Each check is fixed by the mathematics of the scheme, not by its
formulae: closed forms for uniform states, conservation of the distributed
residual (boundary trapezoid integrals of the fluxes and the source),
corner differences that isolate the Lax-Friedrichs term and its dx/dy
weights, linearity in diff, and a linear pressure field at rest.

The reference's cell residual (:673-678) takes the x flux difference as
+(right - left) and the y flux difference as -(top - bottom), neither scaled
by an edge length, against a source scaled by dx dy; the checks below hold
that scheme as written (which is also why it has no x-y transposition
symmetry to test)."""
import numpy as np
import pytest

import pyoracle as po

GAMMA = po.GAMMA_REF


def energy_source(rho, vx, vy):
    """Minus the energy source of get_source (benchmark_rd.f90:747-767) as the
    cell residual calls it: compute_update_rd hands it the primitive vector
    (rho, vx, vy, p) (:681-684) and get_source converts that to primitive
    variables again (:756), so its velocities are vx/rho and vy/rho and the
    energy source is -rho (vx/rho + vy/rho), not -rho (vx + vy)."""
    return rho * (vx / rho + vy / rho)


def uniform(nx, ny, rho, a, b, p):
    u1 = po.rd_conservative(GAMMA, [rho, a, b, p])
    return np.asfortranarray(np.broadcast_to(u1[:, None, None], (4, nx, ny)).copy())


def random_state(nx, ny, seed):
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.3 * rng.random((nx, ny))
    vx = 0.4 * rng.standard_normal((nx, ny))
    vy = 0.4 * rng.standard_normal((nx, ny))
    p = 1.0 + 0.3 * rng.random((nx, ny))
    u = np.zeros((4, nx, ny), order="F")
    u[0] = rho
    u[1] = rho * vx
    u[2] = rho * vy
    u[3] = p / (GAMMA - 1.0) + 0.5 * rho * (vx ** 2 + vy ** 2)
    return u


@pytest.mark.parametrize("state", [(1.3, 0.4, -0.25, 0.8), (2.0, 1.0, 1.0, 1.0)])
@pytest.mark.parametrize("bc", [po.RD_BC_NONE, po.RD_BC_TRANSMISSIVE, po.RD_BC_PERIODIC])
def test_uniform_state_closed_form(bc, state):
    """For a uniform state every flux difference vanishes exactly and u_k =
    ubar, so a cell's corners each receive dt*(-s dx dy)/4: interior nodes get
    dt dx dy (0, rho, rho, rho (a/rho + b/rho)), edges half, corners a quarter,
    before the boundary copies (benchmark_rd.f90:722-737) are applied.  At
    rho = 2, v = (1, 1) the energy entry is 2 dt dx dy; a source that skips
    get_source's second primitive conversion gives 4."""
    nx, ny = 9, 7
    rho, a, b, pr = state
    p = po.make_rd_params(nx, ny, bc=bc, boxlen_x=1.0, boxlen_y=0.75)
    dx, dy, dt = 1.0 / (nx - 1), 0.75 / (ny - 1), 1e-3
    d = po.rd_update(p, uniform(nx, ny, rho, a, b, pr), np.zeros((4, nx, ny), order="F"), dt)
    full = dt * dx * dy * np.array([0.0, rho, rho, energy_source(rho, a, b)])
    count = np.full((nx, ny), 4.0)
    count[0, :] /= 2; count[-1, :] /= 2; count[:, 0] /= 2; count[:, -1] /= 2
    want = full[:, None, None] * count[None] / 4.0
    if bc == po.RD_BC_TRANSMISSIVE:
        want = np.broadcast_to(full[:, None, None], want.shape)
    elif bc == po.RD_BC_PERIODIC:
        want = want.copy()
        want[:, 0, :] = want[:, -1, :]          # x copies first, then y
        want[:, :, -1] = want[:, :, 0]
    np.testing.assert_allclose(d, want, rtol=1e-14, atol=1e-22)


def test_distributed_residual_is_conservative():
    """With dx = dy the Lax-Friedrichs parts of a cell sum to zero, so the sum
    of dudt over all nodes (no boundary copies) equals dx dy/4 times the
    cell-count-weighted sum of diff plus dt times the telescoped cell
    residuals: the trapezoid sums of f on the right minus the left boundary,
    minus those of g on the top minus the bottom, minus dx dy times the
    trapezoid sum of the source."""
    n = 12
    p = po.make_rd_params(n, n, bc=po.RD_BC_NONE)
    h, dt = 1.0 / (n - 1), 3e-3
    u = random_state(n, n, 4)
    diff = 1e-3 * np.asfortranarray(np.random.default_rng(5).standard_normal((4, n, n)))
    d = po.rd_update(p, u, diff, dt)
    rho = u[0]; vx = u[1] / rho; vy = u[2] / rho
    pr = (GAMMA - 1.0) * (u[3] - 0.5 * rho * (vx ** 2 + vy ** 2))
    f = np.stack([rho * vx, rho * vx * vx + pr, rho * vx * vy, vx * (u[3] + pr)])
    g = np.stack([rho * vy, rho * vx * vy, rho * vy * vy + pr, vy * (u[3] + pr)])
    s = np.stack([0 * rho, -rho, -rho, -energy_source(rho, vx, vy)])
    tw = np.ones(n); tw[0] = tw[-1] = 0.5                     # trapezoid weights
    fx = np.einsum("vj,j->v", f[:, -1, :], tw) - np.einsum("vj,j->v", f[:, 0, :], tw)
    gy = np.einsum("vi,i->v", g[:, :, -1], tw) - np.einsum("vi,i->v", g[:, :, 0], tw)
    cnt = np.outer(tw, tw) * 4
    src = h * h * np.einsum("vij,ij->v", s, cnt) / 4
    want = h * h / 4 * np.einsum("vij,ij->v", diff, cnt) + dt * (fx - gy - src)
    np.testing.assert_allclose(d.sum(axis=(1, 2)), want, rtol=1e-11, atol=1e-15)


def quirk_speed(u):
    """The Lax-Friedrichs coefficient's speed as benchmark_rd.f90:692-695
    computes it: compute_speed receives the primitive vector and converts it
    once more."""
    w = np.array([u[0], u[1] / u[0], u[2] / u[0], 0.0])
    w[3] = (GAMMA - 1.0) * (u[3] - 0.5 * w[0] * (w[1] ** 2 + w[2] ** 2))
    w2 = np.array([w[0], w[1] / w[0], w[2] / w[0], 0.0])
    w2[3] = (GAMMA - 1.0) * (w[3] - 0.5 * w2[0] * (w2[1] ** 2 + w2[2] ** 2))
    return np.sqrt(w2[1] ** 2 + w2[2] ** 2) + np.sqrt(GAMMA * max(w2[3], 1e-10) / max(w2[0], 1e-10))


def test_corner_differences_isolate_the_lax_friedrichs_term():
    """phi_11 - phi_12 = dt alpha dy (u_11 - u_12) and phi_21 - phi_22 =
    dt alpha dx (u_21 - u_22) (diff = 0): the common total residual cancels,
    what remains fixes the node of each corner and which spacing weights it
    (dx != dy here)."""
    nx, ny = 6, 5
    p = po.make_rd_params(nx, ny, boxlen_x=1.0, boxlen_y=0.6)
    dx, dy, dt = 1.0 / (nx - 1), 0.6 / (ny - 1), 2e-3
    u = random_state(nx, ny, 11)
    z = np.zeros((4, nx, ny), order="F")
    for (i, j) in [(0, 0), (2, 1), (4, 3)]:
        phi = po.rd_cell(p, u, z, dt, i, j)
        alpha = max(quirk_speed(u[:, a, b]) for a, b in [(i, j), (i + 1, j), (i, j + 1), (i + 1, j + 1)])
        np.testing.assert_allclose(phi[0] - phi[1], dt * alpha * dy * (u[:, i, j] - u[:, i + 1, j]),
                                   rtol=1e-9, atol=1e-16)
        np.testing.assert_allclose(phi[2] - phi[3], dt * alpha * dx * (u[:, i, j + 1] - u[:, i + 1, j + 1]),
                                   rtol=1e-9, atol=1e-16)


def test_linear_in_diff():
    """dudt(u, diff) - dudt(u, 0) = dx dy/4 (adjacent cells) diff at every node."""
    nx, ny = 8, 6
    p = po.make_rd_params(nx, ny, bc=po.RD_BC_NONE, boxlen_y=0.5)
    dx, dy = 1.0 / (nx - 1), 0.5 / (ny - 1)
    u = random_state(nx, ny, 2)
    diff = np.asfortranarray(np.random.default_rng(3).standard_normal((4, nx, ny)))
    d1 = po.rd_update(p, u, diff, 1e-3) - po.rd_update(p, u, np.zeros_like(diff), 1e-3)
    cnt = np.full((nx, ny), 4.0)
    cnt[0, :] /= 2; cnt[-1, :] /= 2; cnt[:, 0] /= 2; cnt[:, -1] /= 2
    np.testing.assert_allclose(d1, dx * dy / 4 * cnt[None] * diff, rtol=1e-10, atol=1e-13)


def test_linear_pressure_at_rest_closed_form():
    """rho constant, v = 0, p = p0 + a x + b y: every cell residual is exactly
    (0, a h + rho h^2, -b h + rho h^2, 0) (flux differences of a linear field,
    the source -rho of both momenta), the momenta carry no Lax-Friedrichs part,
    so each interior node receives dt times that residual."""
    n, rho, p0, a, b, dt = 10, 1.4, 2.0, 0.3, -0.2, 1e-3
    p = po.make_rd_params(n, n, bc=po.RD_BC_NONE)
    h = 1.0 / (n - 1)
    x = np.linspace(0.0, 1.0, n)[:, None]
    y = np.linspace(0.0, 1.0, n)[None, :]
    u = np.zeros((4, n, n), order="F")
    u[0] = rho
    u[3] = (p0 + a * x + b * y) / (GAMMA - 1.0)
    d = po.rd_update(p, u, np.zeros_like(u), dt)
    np.testing.assert_array_equal(d[0], 0.0)
    np.testing.assert_allclose(d[1, 1:-1, 1:-1], dt * (a * h + rho * h * h), rtol=1e-10)
    np.testing.assert_allclose(d[2, 1:-1, 1:-1], dt * (-b * h + rho * h * h), rtol=1e-10)


def test_max_speed_and_step_closed_forms():
    """compute_max_speed of a uniform state is |v| + sqrt(gamma p/rho); one
    'RD1' step takes dt = cfl dx dy/cmax/3 and, with transmissive copies, moves
    the uniform state by -dt (0, rho, rho, rho (a/rho + b/rho))."""
    nx, ny, rho, a, b, pr = 7, 9, 0.9, 0.3, 0.5, 1.1
    p = po.make_rd_params(nx, ny, bc=po.RD_BC_TRANSMISSIVE, cfl=0.4)
    dx, dy = 1.0 / (nx - 1), 1.0 / (ny - 1)
    u = uniform(nx, ny, rho, a, b, pr)
    c = po.rd_max_speed(p, u)
    assert c == pytest.approx(np.hypot(a, b) + np.sqrt(GAMMA * pr / rho), rel=1e-15)
    u0 = u.copy(order="F")
    dt = po.rd_step(p, u)
    assert dt == pytest.approx(0.4 * dx * dy / c / 3.0, rel=1e-15)
    want = u0 - dt * np.array([0.0, rho, rho, energy_source(rho, a, b)])[:, None, None]
    np.testing.assert_allclose(u, want, rtol=1e-13, atol=1e-16)


def test_initial_conditions():
    """Cases 1, 2 and 8 of get_initial_conditions on the node grid
    x_i = (i-1) boxlen_x/(nx-1), through compute_conservative."""
    nx, ny = 11, 6
    p = po.make_rd_params(nx, ny)
    x = np.linspace(0.0, 1.0, nx)[:, None]
    y = np.linspace(0.0, 1.0, ny)[None, :]
    r121 = float(np.float32(1.21))
    u2 = po.rd_initial(p, 2)
    np.testing.assert_allclose(u2[0], r121 * np.exp(-r121 * (x + y)), rtol=1e-15)
    np.testing.assert_array_equal(u2[1:3], 0.0)
    np.testing.assert_allclose(u2[3], np.exp(-r121 * (x + y)) / (GAMMA - 1.0), rtol=1e-15)
    u1 = po.rd_initial(p, 1)
    rho1 = 1.0 + np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2) * 20)
    np.testing.assert_allclose(u1[0], rho1, rtol=1e-15)
    np.testing.assert_allclose(u1[1], rho1, rtol=1e-15)
    eta = 1e-2
    u8 = po.rd_initial(p, 8, eta)
    k8 = float(np.float32(100.0) * np.float32(1.21))
    p8 = np.exp(-r121 * (x + y)) + eta * np.exp(-k8 * ((x - float(np.float32(0.3))) ** 2 +
                                                        (y - float(np.float32(0.3))) ** 2))
    np.testing.assert_allclose(u8[3], p8 / (GAMMA - 1.0), rtol=1e-14)
    with pytest.raises(ValueError):
        po.rd_initial(p, 5)


def test_bad_parameters():
    z = np.zeros((4, 4, 4), order="F")
    with pytest.raises(ValueError):
        po.rd_update(po.make_rd_params(4, 4, bc=1), z, z, 1e-3)
    with pytest.raises(ValueError):
        po.rd_update(po.make_rd_params(1, 4), np.zeros((4, 1, 4), order="F"), np.zeros((4, 1, 4), order="F"), 1e-3)
