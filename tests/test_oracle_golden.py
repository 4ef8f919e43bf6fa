"""The CPU oracle against the golden fixtures written from the flang-built
reference (tests/golden/make_golden.py), and live against oracle/_ref when it
is present.  The oracle is bit-identical to the reference in this container;
the tolerances allow a few ulp for a libm exp that differs on another CPU."""
import os

import numpy as np
import pytest

import pyoracle as po
from make_golden import bench2d_inputs, summarize
import ref_runner as rr

GOLD = os.path.dirname(os.path.abspath(__file__)) + "/golden"
EPS = np.finfo(np.float64).eps


def load(name):
    with np.load(os.path.join(GOLD, name + ".npz")) as z:
        return {k: z[k] for k in z.files}


def flux_scale(nx):
    # |dudt| errors from a few-ulp change in O(1) face fluxes divided by dx
    return 16 * EPS * 2.0 * nx


def test_update_exact_matches_reference_golden():
    g = load("bench2d_update_exact_300")
    p, u, w = bench2d_inputs(salt=int(g["salt"]))
    d = po.ref2d_update_exact(p, 2, u, w)
    s = summarize(d)
    np.testing.assert_array_equal(s["idx"], g["idx"])
    tol = flux_scale(300)
    assert np.max(np.abs(s["sample"] - g["sample"])) <= tol
    assert np.max(np.abs(s["maxabs"] - g["maxabs"])) <= tol
    np.testing.assert_allclose(s["sum"], g["sum"], rtol=0, atol=tol * 300 * 300)
    np.testing.assert_allclose(s["abssum"], g["abssum"], rtol=1e-12)


def test_update_plain_matches_reference_golden_bitwise():
    """compute_update (benchmark_2d.f90:370-463), the plain non-well-balanced
    residual: no exp on this path, so the restatement is bit-identical."""
    g = load("bench2d_update_plain_300")
    p, u, w = bench2d_inputs(salt=int(g["salt"]), amp_bump=float(g["amp_bump"]), amp_noise=float(g["amp_noise"]))
    s = summarize(po.ref2d_update_plain(p, u))
    for k in ("idx", "sample", "sum", "abssum", "maxabs"):
        np.testing.assert_array_equal(s[k], g[k])


def test_update_plain_is_not_well_balanced():
    """The plain residual of a resting atmosphere is a truncation error: the
    Rusanov diffusion of the stratified density is first order, the imbalance of
    pressure differences against the pointwise source second order."""
    errs = []
    for n in (40, 80):
        p = po.make_params(n, n, gamma=po.GAMMA_REF)
        w = po.equilibrium_primitive(p)
        d = po.ref2d_update_plain(p, po.conservative(p.gamma, w))
        errs.append((np.max(np.abs(d[0])), np.max(np.abs(d[2]))))
    assert errs[0][0] > 1e-2 and 1.6 < errs[0][0] / errs[1][0] < 2.4
    assert errs[0][1] > 1e-5 and 3.2 < errs[0][1] / errs[1][1] < 4.8


def test_update_exact_equilibrium_is_zero_like_reference():
    g = load("bench2d_update_exact_eq_300")
    assert int(g["nonzero"]) == 0 and np.all(g["maxabs"] == 0.0)
    p, u, w = bench2d_inputs()
    ueq = po.conservative(p.gamma, w)
    assert np.count_nonzero(po.ref2d_update_exact(p, 2, ueq, w)) == 0


def test_max_speed_matches_reference_golden():
    g = load("bench2d_max_speed_300")
    p, u, w = bench2d_inputs(salt=int(g["salt"]))
    assert abs(po.ref2d_max_speed(p, u) - float(g["cmax"][0])) <= 4 * EPS


def test_evolve_matches_reference_golden():
    g = load("bench2d_evolve_300")
    p, u, w = bench2d_inputs(salt=int(g["salt"]))
    ue, t, n = po.ref2d_evolve(p, 2, u, w, float(g["tend"]))
    assert n == 3
    s = summarize(ue - po.conservative(p.gamma, w))
    np.testing.assert_array_equal(s["idx"], g["idx"])
    assert np.max(np.abs(s["sample"] - g["sample"])) <= 1e-13
    np.testing.assert_allclose(s["abssum"], g["abssum"], rtol=1e-11)


def test_update_exact_speed_and_evolve_with_dx_ne_dy_match_reference_golden():
    """The reference's 2D program compiled at nx = 300, ny = 200 (a sed edit of
    parameters_2d.f90, oracle/Makefile): with dx != dy, compute_update_exact
    places the y faces at (j-1)*dx (benchmark_2d.f90:513) and evolve takes dt
    from dx alone (:242); the restatement follows both."""
    g = load("bench2d_update_exact_300x200")
    p, u, w = bench2d_inputs(salt=int(g["salt"]), nx=300, ny=200)
    d = po.ref2d_update_exact(p, 2, u, w)
    s = summarize(d)
    np.testing.assert_array_equal(s["idx"], g["idx"])
    tol = flux_scale(300)
    assert np.max(np.abs(s["sample"] - g["sample"])) <= tol
    assert np.max(np.abs(s["maxabs"] - g["maxabs"])) <= tol
    np.testing.assert_allclose(s["abssum"], g["abssum"], rtol=1e-12)
    assert abs(po.ref2d_max_speed(p, u) - float(g["cmax"][0])) <= 4 * EPS
    # a restatement with the y faces at (j-1)*dy would differ at O(1) in the
    # y fluxes: the golden must tell them apart
    ge = load("bench2d_evolve_300x200")
    ue, t, n = po.ref2d_evolve(p, 2, u, w, float(ge["tend"]))
    assert n == 3
    se = summarize(ue - po.conservative(p.gamma, w))
    np.testing.assert_array_equal(se["idx"], ge["idx"])
    assert np.max(np.abs(se["sample"] - ge["sample"])) <= 1e-13
    np.testing.assert_allclose(se["abssum"], ge["abssum"], rtol=1e-11)


def test_minmod_matches_reference_golden():
    g = load("minmod")
    mm = np.array([po.minmod(*t) for t in g["xyz"]])
    gm = np.array([po.generalized_minmod(*t, float(g["M"]), float(g["dx"])) for t in g["xyz"]])
    np.testing.assert_array_equal(mm, g["minmod"])
    np.testing.assert_array_equal(gm, g["generalized_minmod"])


def test_hllflux_as_written_matches_reference_golden():
    g = load("dg2d_hllflux")
    out = np.array([po.ref_hllflux(float(g["gamma"]), a, b, c, d)
                    for a, b, c, d in zip(g["ul"], g["ur"], g["fl"], g["fr"])])
    np.testing.assert_allclose(out, g["f"], rtol=1e-14, atol=1e-14)


def test_riemann_hllc_matches_reference_golden_bitwise():
    """riemann_hllc (dg_with_source.f90:1318-1374) restated statement for
    statement: equal to the flang build's output bit for bit on states that reach
    all four sampled regions."""
    g = load("dg1d_riemann_hllc")
    gam = float(g["gamma"])
    out = np.array([po.ref_riemann_hllc_1d(gam, a, b) for a, b in zip(g["ul"], g["ur"])])
    np.testing.assert_array_equal(out, g["f"])
    # the fixture covers every region: supersonic both ways and both star states
    regions = set()
    for a, b in zip(g["ul"], g["ur"]):
        wl, wr = a[1] / a[0], b[1] / b[0]
        pl = (gam - 1) * (a[2] - 0.5 * a[0] * wl * wl)
        pr = (gam - 1) * (b[2] - 0.5 * b[0] * wr * wr)
        cl, cr = np.sqrt(gam * pl / a[0]), np.sqrt(gam * pr / b[0])
        sl, sr = min(wl, wr) - max(cl, cr), max(wl, wr) + max(cl, cr)
        regions.add("L" if sl > 0 else ("R" if sr <= 0 else "star"))
    assert regions == {"L", "R", "star"}


def test_hllc_2d_reduces_to_reference_riemann_hllc():
    """The 2D HLLC of the configured schemes, on states without tangential
    velocity, along x and along y (rotated), against the reference's own 1D HLLC:
    the same fluxes to rounding (its star pressure is rearranged), zero tangential
    flux."""
    g = load("dg1d_riemann_hllc")
    gam = float(g["gamma"])
    for a, b, f in zip(g["ul"], g["ur"], g["f"]):
        for d in (0, 1):
            ua = np.zeros(4); ub = np.zeros(4)
            ua[0], ua[1 + d], ua[3] = a
            ub[0], ub[1 + d], ub[3] = b
            f2 = po.hllc_2d(gam, d, ua, ub)
            want = np.array([f[0], 0.0, 0.0, f[2]])
            want[1 + d] = f[1]
            np.testing.assert_allclose(f2, want, rtol=1e-12, atol=1e-12 * np.abs(f).max())


@pytest.mark.skipif(not rr.available("bench2d"), reason="oracle/_ref not built")
def test_update_exact_live_against_reference_other_input():
    """A second input (other noise salt, larger bump) straight from the flang
    build: the restatement is bit-identical."""
    p, u, w = bench2d_inputs(salt=5, amp_bump=1e-2, amp_noise=1e-3)
    ref = rr.run("bench2d_update_exact", u=u, w_eq=w)["dudt"]
    mine = po.ref2d_update_exact(p, 2, u, w)
    assert np.max(np.abs(ref - mine)) <= flux_scale(300)


@pytest.mark.skipif(not rr.available("bench2d"), reason="oracle/_ref not built")
def test_equilibrium_nequilibrium_cases_live():
    """get_equilibrium_solution cases 1 and 2 (benchmark_2d.f90:190-202) through
    the configured closed form with (rho0,p0,g) = (1,1,1) / (1.21f,1,1).  The
    reference module fixes nequilibrium=2; case 1 is the same formula with
    rho_0 = p_0 = g = 1."""
    x, y = np.meshgrid(np.linspace(0, 1, 7), np.linspace(0, 1, 5), indexing="ij")
    ref = rr.run("bench2d_equilibrium", x=np.asfortranarray(x), y=np.asfortranarray(y))["w"]
    p = po.make_params(8, 8)
    k = p.rho0 * p.g / p.p0
    mine = np.zeros_like(ref)
    mine[0] = p.rho0 * np.exp(-(k * (x + y)))
    mine[3] = p.p0 * np.exp(-(k * (x + y)))
    np.testing.assert_allclose(mine, ref, rtol=4 * EPS, atol=0)
