"""The reference's own 2D DG program on the GPU (fvwb_dgq_*, FP64) against the
fixtures of the flang builds of 2d/benchmark_2d_dg.f90 (the shipped set-up and
three edits of its switches, tests/golden/dgq_*.npz) and against the oracle on
grids the fixtures do not cover.  Every comparison is bit for bit except the
device's compute_error, whose parallel sums round differently from the
reference's serial loop (relative 1e-13)."""
import os
import sys

import numpy as np
import pytest

import fvwb
import pyoracle as po
from make_golden import DGQ_BUILDS

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = {"dg2d": "shipped", "dgq_hll2": "hll2", "dgq_llf1": "llf1", "dgq_hllc": "hllc"}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def load(lib):
    with np.load(os.path.join(GOLD, f"dgq_{NAMES[lib]}.npz")) as z:
        return {k: z[k] for k in z.files}


def gparams(n, kw):
    return fvwb.reference_params_dgq(n, **kw)


@pytest.mark.parametrize("lib", list(DGQ_BUILDS))
def test_dropins_equal_reference_build(lib):
    say(f"case: dgq drop-ins vs the reference build {lib}")
    g = load(lib)
    with fvwb.SolverDGQ(gparams(8, DGQ_BUILDS[lib])) as s:
        np.testing.assert_array_equal(s.compute_update(g["delta_u"], g["x"], g["y"]), g["dudt"])
        if "dudt_x6" in g:
            np.testing.assert_array_equal(s.compute_update(g["delta_u"], 6 * g["x"], 6 * g["y"]), g["dudt_x6"])
        ms = s.compute_max_speed(np.asfortranarray(g["delta_u"][:, :, :, 0, 0]))
        np.testing.assert_array_equal(np.array(ms), g["max_speed"])
        np.testing.assert_array_equal(s.apply_limiter(g["limiter_in"]), g["limiter_out"])


def test_projections_equal_reference_build():
    say("case: dgq projections vs get_modes_from_nodes / get_nodes_from_modes")
    with np.load(os.path.join(GOLD, "dg2d_projections.npz")) as z:
        a, m, n = z["a"], z["modes_from_nodes"], z["nodes_from_modes"]
    with fvwb.SolverDGQ(fvwb.reference_params_dgq()) as s:
        np.testing.assert_array_equal(s.modes_from_nodes(a), m)
        np.testing.assert_array_equal(s.nodes_from_modes(a), n)


@pytest.mark.parametrize("lib", list(DGQ_BUILDS))
def test_resident_evolve_equals_reference_run(lib):
    """evolve from the build's own initial condition on the device: the
    projection and the first apply_limiter through the drop-ins, then the
    resident loop to tend; the step count, t, the final nodal state and
    compute_error's norms equal the reference's run."""
    g = load(lib)
    say(f"case: dgq resident evolve vs the reference run {lib}, tend {float(g['tend'])}")
    with fvwb.SolverDGQ(gparams(8, DGQ_BUILDS[lib])) as s:
        du = s.apply_limiter(s.modes_from_nodes(g["u0"]))
        s.upload(du, g["x"], g["y"])
        s.advance(t_end=float(g["tend"]))
        t, dt_last, n = s.time()
        nodes = s.nodes_from_modes(s.download())
        linf, l1, l2 = s.compute_error(g["u0"])
        linf2, l1b, _ = s.compute_error(g["u0"], nodes)
    assert n == len(g["evolve_dt"]) and t == float(g["tend"]) and dt_last == g["evolve_dt"][-1]
    np.testing.assert_array_equal(nodes, g["evolve_u"])
    printed = dict(zip([str(k) for k in g["norm_keys"]], g["norms"]))
    for v, nm in enumerate(("rho", "velx", "vely", "energy")):
        assert linf[v] == printed[f"maxerror {nm}"]
        np.testing.assert_allclose(l1[v], printed[f"l1error {nm}"], rtol=1e-13)
        np.testing.assert_allclose(l2[v], printed[f"l2error {nm}"], rtol=1e-13)
    np.testing.assert_array_equal(linf2, linf)
    np.testing.assert_array_equal(l1b, l1)


@pytest.mark.parametrize("order,kw", [
    (3, dict(flux_type=fvwb.DGQ_FLUX_HLL2, source=2, grad_phi_case=2, bc=1, solver=fvwb.DGQ_SOLVER_RK4)),
    (2, dict(flux_type=fvwb.DGQ_FLUX_HLLC, source=2, grad_phi_case=1, bc=2, solver=fvwb.DGQ_SOLVER_SS4)),
    (3, dict(flux_type=fvwb.DGQ_FLUX_LLF1, source=3, bc=3, ninit=12, solver=fvwb.DGQ_SOLVER_EQL, boxlen_x=6.0,
             boxlen_y=6.0)),
])
def test_larger_grids_equal_oracle(order, kw):
    """48 x 48 cells (past the 8 x 8 the reference is compiled at): residual and
    five resident steps against the oracle, bit for bit."""
    n = 48
    say(f"case: dgq {n}x{n} order {order} {kw} vs oracle")
    okw = dict(kw)
    op = po.make_dgq_params(n, order=order, **okw)
    gp = gparams(n, dict(order=order, **kw))
    x, y = po.dgq_coords(op)
    rng = np.random.default_rng(3)
    nodes = np.zeros((4, n, n, order, order), order="F")
    nodes[0] = 1.0 + 0.3 * rng.random(nodes[0].shape)
    nodes[1] = 0.3 * rng.standard_normal(nodes[0].shape)
    nodes[2] = 0.3 * rng.standard_normal(nodes[0].shape)
    nodes[3] = 2.0 + 0.5 * rng.random(nodes[0].shape)
    with fvwb.SolverDGQ(gp) as s:
        du = s.modes_from_nodes(nodes)
        ref = np.zeros_like(du)
        po.lib().orc_ref_dg2d_modes_from_nodes(n, n, order, po._p(nodes), po._p(ref))
        np.testing.assert_array_equal(du, ref)
        np.testing.assert_array_equal(s.compute_update(du, x, y), po.refdg_update(op, du, x, y))
        dl = s.apply_limiter(du)
        np.testing.assert_array_equal(dl, po.refdg_limiter(op, du))
        s.upload(dl, x, y)
        s.advance(5)
        out, t, k, dts = po.refdg_evolve(op, dl, x, y, float("inf"), max_steps=5)
        tg, dtg, kg = s.time()
        assert kg == k == 5 and tg == t and dtg == dts[-1]
        np.testing.assert_array_equal(s.download(), out)


def test_advance_before_upload_and_bad_shapes():
    with fvwb.SolverDGQ(fvwb.reference_params_dgq()) as s:
        with pytest.raises(fvwb.FvwbError) as e:
            s.advance(1)
        assert e.value.code == fvwb.ERR_STATE
        with pytest.raises(fvwb.FvwbError):
            s.compute_update(np.zeros((4, 8, 8, 3, 3)), np.zeros((8, 8, 2, 2)), np.zeros((8, 8, 2, 2)))
