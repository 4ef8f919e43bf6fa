"""The oracle of the reference's own 2D DG program (oracle/fvwb_oracle_dgq.c)
against the flang builds of 2d/benchmark_2d_dg.f90: the shipped set-up (8x8
cells, mx = my = 2, flux_type 'llf' (no branch: zero face flux), source 1,
'RK4', 'ONP') and three edits of its compile-time switches (oracle/Makefile):

  hll2: mx = my = 3, 'hll2', source 2 with grad_phi_case 2 (point mass), 'DEB'
  llf1: 'llf1', source 2 with grad_phi_case 1, bc 2, ninit 2, 'SS4'
  hllc: mx = my = 3, 'hllc', source 3, bc 3, ninit 12, 'EQL'

Fixtures (tests/golden/dgq_*.npz, written by make_golden.py): the build's own
coordinates and initial state, compute_update on a perturbed state,
compute_max_speed of its means, apply_limiter on steep modes, and a whole
evolve run (every dt from its 'time=' lines, the final nodal state,
compute_error's printed norms, the snapshot files of output_file).  Every
comparison is bit for bit except the printed norms (17 significant digits)."""
import os

import numpy as np
import pytest

import pyoracle as po
import ref_runner as rr
from make_golden import DGQ_BUILDS

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = {"dg2d": "shipped", "dgq_hll2": "hll2", "dgq_llf1": "llf1", "dgq_hllc": "hllc"}


def load(lib):
    with np.load(os.path.join(GOLD, f"dgq_{NAMES[lib]}.npz")) as z:
        return {k: z[k] for k in z.files}


def params(lib):
    return po.make_dgq_params(8, **DGQ_BUILDS[lib])


def modes_of(p, nodes):
    m = np.zeros_like(nodes, order="F")
    po.lib().orc_ref_dg2d_modes_from_nodes(p.nx, p.ny, p.order, po._p(np.asfortranarray(nodes)), po._p(m))
    return m


def nodes_of(p, modes):
    m = np.zeros_like(modes, order="F")
    po.lib().orc_ref_dg2d_nodes_from_modes(p.nx, p.ny, p.order, po._p(np.asfortranarray(modes)), po._p(m))
    return m


def fmt_snapshot(p, nodes):
    """output_file (2d/benchmark_2d_dg.f90:468-495) with var = 1: per cell, x
    outer and y inner, the first node's coordinates and the primitive state
    minus the equilibrium (nequilibrium 3 of the shipped set-up: zero) in
    '(7(1PE12.5,1X))'."""
    x, y = po.dgq_coords(p)
    lines = []
    for i in range(p.nx):
        for j in range(p.ny):
            u = nodes[:, i, j, 0, 0]
            r = max(u[0], float(np.float32(10e-10)))
            vx, vy = u[1] / r, u[2] / r
            pr = (p.gamma - 1.0) * (u[3] - 0.5 * r * (vx * vx + vy * vy))
            vals = [x[i, j, 0, 0], y[i, j, 0, 0], r, vx, vy, pr]
            lines.append(" ".join(f"{v:12.5E}" for v in vals))    # the format stops after the last item
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("lib", list(DGQ_BUILDS))
def test_coords_match_get_coords(lib):
    g, p = load(lib), params(lib)
    x, y = po.dgq_coords(p)
    np.testing.assert_array_equal(x, g["x"])
    np.testing.assert_array_equal(y, g["y"])


@pytest.mark.parametrize("lib", list(DGQ_BUILDS))
def test_compute_update_bitwise(lib):
    g, p = load(lib), params(lib)
    d = po.refdg_update(p, g["delta_u"], g["x"], g["y"])
    assert np.abs(g["dudt"]).max() > 1.0
    np.testing.assert_array_equal(d, g["dudt"])
    if "dudt_x6" in g:            # ninit 12: special_boundary_conditions freezes r > 2
        d6 = po.refdg_update(p, g["delta_u"], 6 * g["x"], 6 * g["y"])
        assert np.count_nonzero(g["dudt_x6"] == 0) > 0
        np.testing.assert_array_equal(d6, g["dudt_x6"])


def test_shipped_flux_type_matches_no_branch():
    """flux_type 'llf' selects no branch of compute_num_flux, so the shipped
    residual has no surface term.  For a spatially constant state only the
    volume integrals of the linear modes survive: dudt(1,0) = (1/dx) F_x
    sqrt(3) (sum_a w_a)(sum_b w_b) / 2 = 16 sqrt(3) F_x (dx = 1/8), likewise
    dudt(0,1) with F_y, and the mean and (1,1) modes stay exactly zero.  With
    'llf1' and periodic walls the face integrals cancel the volume ones: the
    constant state is at rest to rounding."""
    p = params("dg2d")
    x, y = po.dgq_coords(p)
    du = np.zeros((4, 8, 8, 2, 2), order="F")
    u = np.array([1.3, 0.2, -0.1, 2.5])
    du[:, :, :, 0, 0] = u[:, None, None]
    d = po.refdg_update(p, du, x, y)
    g = p.gamma
    vx, vy = u[1] / u[0], u[2] / u[0]
    pr = (g - 1) * (u[3] - 0.5 * u[0] * (vx * vx + vy * vy))
    fx = np.array([u[1], u[1] * vx + pr, u[1] * vy, vx * (u[3] + pr)])
    fy = np.array([u[2], u[2] * vx, u[2] * vy + pr, vy * (u[3] + pr)])
    assert np.count_nonzero(d[:, :, :, 0, 0]) == 0 and np.count_nonzero(d[:, :, :, 1, 1]) == 0
    np.testing.assert_allclose(d[:, :, :, 1, 0], np.broadcast_to((16 * np.sqrt(3) * fx)[:, None, None], (4, 8, 8)), rtol=1e-14)
    np.testing.assert_allclose(d[:, :, :, 0, 1], np.broadcast_to((16 * np.sqrt(3) * fy)[:, None, None], (4, 8, 8)), rtol=1e-14)
    q = po.make_dgq_params(8, flux_type=po.DGQ_FLUX_LLF1, bc=1)
    assert np.abs(po.refdg_update(q, du, x, y)).max() <= 1e-13


@pytest.mark.parametrize("lib", list(DGQ_BUILDS))
def test_max_speed_and_limiter_bitwise(lib):
    g, p = load(lib), params(lib)
    ms = po.refdg_max_speed(p, np.asfortranarray(g["delta_u"][:, :, :, 0, 0]))
    np.testing.assert_array_equal(np.array(ms), g["max_speed"])
    out = po.refdg_limiter(p, g["limiter_in"])
    assert np.count_nonzero(out != g["limiter_in"]) > 0
    np.testing.assert_array_equal(out, g["limiter_out"])


@pytest.mark.parametrize("lib", list(DGQ_BUILDS))
def test_evolve_bitwise(lib):
    """evolve from the build's own initial condition: get_modes_from_nodes,
    apply_limiter, then the loop: every dt, the final nodal state, and
    compute_error's norms against the initial condition."""
    g, p = load(lib), params(lib)
    du = po.refdg_limiter(p, modes_of(p, g["u0"]))
    out, t, n, dts = po.refdg_evolve(p, du, g["x"], g["y"], float(g["tend"]))
    assert n == len(g["evolve_dt"]) and t == float(g["tend"])
    np.testing.assert_array_equal(dts, g["evolve_dt"])
    np.testing.assert_array_equal(np.cumsum(dts)[-1], g["evolve_t"][-1])
    nodes = nodes_of(p, out)
    np.testing.assert_array_equal(nodes, g["evolve_u"])
    linf, l1, l2 = po.refdg_error(p, nodes, g["u0"])
    printed = dict(zip([str(k) for k in g["norm_keys"]], g["norms"]))
    names = ("rho", "velx", "vely", "energy")
    for v, nm in enumerate(names):
        assert linf[v] == pytest.approx(printed[f"maxerror {nm}"], rel=1e-15, abs=1e-300)
        assert l1[v] == pytest.approx(printed[f"l1error {nm}"], rel=1e-15, abs=1e-300)
        assert l2[v] == pytest.approx(printed[f"l2error {nm}"], rel=1e-15, abs=1e-300)


def test_snapshot_cadence_and_format():
    """make_movie with interval 50 (2d/parameters_dg_2d.f90:9,39): the shipped
    run to t = 0.05 (66 steps) writes derp.dat before the loop and SIM00001.dat
    after step 50 (:759-766), each one line per cell in '(7(1PE12.5,1X))'."""
    g, p = load("dg2d"), params("dg2d")
    assert [str(f) for f in g["files"]] == ["SIM00001.dat", "derp.dat"]
    du = po.refdg_limiter(p, modes_of(p, g["u0"]))
    assert str(g["file_derp.dat"]) == fmt_snapshot(p, nodes_of(p, du))
    out, t, n, _ = po.refdg_evolve(p, du, g["x"], g["y"], float(g["tend"]), max_steps=50)
    assert n == 50
    assert str(g["file_SIM00001.dat"]) == fmt_snapshot(p, nodes_of(p, out))


def test_compute_error_closed_forms():
    """compute_error on constructed differences: L_inf is the largest nodal
    difference; a difference d constant on the nodes gives L1 = |d| and L2 = |d|
    (the Gauss weights sum to 2 per direction, times dx dy/4 per cell over
    nx ny cells), each per variable."""
    p = params("dgq_hll2")
    u0 = np.zeros((4, 8, 8, 3, 3), order="F")
    u = u0.copy(order="F")
    d = np.array([0.5, -0.25, 2.0, 1e-3])
    u += d[:, None, None, None, None]
    u[2, 3, 4, 1, 2] += 7.0
    linf, l1, l2 = po.refdg_error(p, u, u0)
    np.testing.assert_allclose(linf, [0.5, 0.25, 9.0, 1e-3], rtol=1e-15)
    xq, wq = po.dg_gl_quadrature(3)
    extra1 = 7.0 * wq[1] * wq[2] / 4 / 64
    np.testing.assert_allclose(l1, np.abs(d) + [0, 0, extra1, 0], rtol=1e-13)
    extra2 = ((9.0 ** 2 - 2.0 ** 2) * wq[1] * wq[2] / 4 / 64)
    np.testing.assert_allclose(l2, np.sqrt(d ** 2 + [0, 0, extra2, 0]), rtol=1e-13)


@pytest.mark.skipif(not all(rr.available(k) for k in DGQ_BUILDS), reason="oracle/_ref not built")
@pytest.mark.parametrize("lib", list(DGQ_BUILDS))
def test_live_reference_update_on_fresh_state(lib):
    """Against the build itself (where oracle/_ref is present) on a state the
    fixtures do not hold."""
    p = params(lib)
    g = load(lib)
    rng = np.random.default_rng(99)
    nodes = nodes_of(p, g["delta_u"])
    nodes[0] *= 1 + 0.02 * rng.standard_normal(nodes[0].shape)
    nodes[3] *= 1 + 0.02 * rng.random(nodes[0].shape)
    du = modes_of(p, nodes)
    d = rr.run("dgq_update", lib=np.array(lib), delta_u=du, x=g["x"], y=g["y"])["dudt"]
    np.testing.assert_array_equal(po.refdg_update(p, du, g["x"], g["y"]), d)
