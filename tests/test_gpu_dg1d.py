"""The reference's 1D modal DG on the GPU (through the C ABI) against the
reference fixtures and the oracle.  The device kernel rounds statement by
statement like the oracle (no contraction, IEEE division and square root) and
takes the libm-evaluated values (face equilibria, Gauss rule, basis) from host
tables computed with the reference's expressions, so every comparison here is
bit for bit."""
import os
import sys

import numpy as np
import pytest

import fvwb
import pyoracle as po

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def load(name):
    with np.load(os.path.join(GOLD, name + ".npz")) as z:
        return {k: z[k] for k in z.files}


def gpu_p(op):
    return fvwb.DG1DParams(3, op.nx, op.n, op.riemann, op.source, 0, op.gamma, op.boxlen)


def same(a, b, what):
    """Bit for bit, NaN where the other side has NaN (the reference's n >= 4
    set-ups, whose Gauss weights carry an extra factor 2n+1, can blow up)."""
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    if not np.array_equal(a, b, equal_nan=True):
        raise AssertionError(f"{what}: {np.count_nonzero(a != b)} entries differ, max {np.max(np.abs(a - b)):.3e}")


def test_update_exact_delta_matches_reference_golden_bitwise():
    say("case: 1D DG compute_update_exact_delta, 128 cells, n=3, HLLC, reference fixture")
    g = load("dg1d_update_exact_delta_128")
    with fvwb.SolverDG1D(fvwb.reference_params_dg1d(128)) as s:
        for k in ("init", "rand", "strong", "zero"):
            same(s.compute_update_exact_delta(g["du_" + k], g["u_eq"]), g["dudt_" + k], k)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("riemann", [po.DG1D_RIEMANN_LLF, po.DG1D_RIEMANN_HLLC])
@pytest.mark.parametrize("source", [1, 2])
def test_update_exact_delta_matches_oracle_bitwise(n, riemann, source):
    nx = 300
    say(f"case: 1D DG residual n={n} riemann={riemann} source={source} nx={nx}")
    op = po.make_dg1d_params(nx, n=n, riemann=riemann, source=source)
    rng = np.random.default_rng(100 * n + 10 * riemann + source)
    du = np.asfortranarray(rng.normal(size=(3, n, nx)) * 0.03)
    ue = po.dg1d_equilibrium_nodes(op)
    with fvwb.SolverDG1D(gpu_p(op)) as s:
        same(s.compute_update_exact_delta(du, ue), po.dg1d_update_exact_delta(op, du, ue), "dudt")


def test_update_exact_delta_large_and_small_grids():
    for nx in (3, 4, 20000):
        say(f"case: 1D DG residual nx={nx}")
        op = po.make_dg1d_params(nx)
        du = np.asfortranarray(np.random.default_rng(nx).normal(size=(3, 3, nx)) * 0.01)
        ue = po.dg1d_equilibrium_nodes(op)
        with fvwb.SolverDG1D(gpu_p(op)) as s:
            same(s.compute_update_exact_delta(du, ue), po.dg1d_update_exact_delta(op, du, ue), f"nx={nx}")


def test_error_paths():
    say("case: 1D DG argument checks")
    for bad in (dict(n=7), dict(n=0), dict(riemann=3), dict(source=0), dict(nx=2)):
        kw = dict(nx=16)
        kw.update(bad)
        nx = kw.pop("nx")
        with pytest.raises(fvwb.FvwbError):
            fvwb.SolverDG1D(gpu_p(po.make_dg1d_params(nx, **kw)))
    with fvwb.SolverDG1D(fvwb.reference_params_dg1d(16)) as s:
        with pytest.raises(fvwb.FvwbError):
            s.compute_update_exact_delta(np.zeros((3, 2, 16)), np.zeros((3, 3, 16)))


# ---- the main program's time loop ('RKi') ----------------------------------------------
def test_run_matches_reference_program_bitwise():
    """The reference's own run (n = 3, nx = 128, HLLC, ninit = 8, tend = 0.2) on
    the device: the same 236 steps, the last t and dt equal to what the
    reference program printed, the final 'error' equal to its own, and the
    modes and u_init equal to the oracle's run, all bit for bit."""
    say("case: 1D DG RKi run of the reference set-up to t = 0.2")
    g = load("dg1d_run_rki_128")
    op = po.make_dg1d_params(128)
    d0, e, u0 = po.dg1d_initial(op)
    do, uo, to, ko, _, _ = po.dg1d_run_rki(op, d0, e, u0, 0.2)
    with fvwb.SolverDG1D(fvwb.reference_params_dg1d(128)) as s:
        s.upload(d0, e, u0)
        s.advance(t_end=0.2)
        t, dtl, k = s.time()
        d, u = s.download()
    say(f"steps {k}, t {t!r}, dt {dtl!r}")
    assert k == 236 and t == g["t"][-1] and dtl == g["dt"][-1]
    same(d, do, "modes")
    same(u, uo, "u_init")
    chsi, _ = po.dg1d_gl_quadrature(3)
    x = (np.arange(1, 129) - 0.5) / 128 + 1 / 128 / 2.0 * chsi[0]
    U = u[:, 0, :]
    pr = (op.gamma - 1.0) * (U[2] - 0.5 * U[0] * ((U[1] / U[0]) * (U[1] / U[0])))
    assert np.max(np.abs(pr - np.exp(-x))) == float(g["error"])


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("riemann", [po.DG1D_RIEMANN_LLF, po.DG1D_RIEMANN_HLLC])
def test_run_matches_oracle_bitwise(n, riemann):
    """30 steps of a pulse of size 1e-2 on 100 cells, every set-up: modes,
    u_init, t and dt equal to the oracle's loop bit for bit."""
    say(f"case: 1D DG RKi run n={n} riemann={riemann} 30 steps")
    op = po.make_dg1d_params(100, n=n, riemann=riemann)
    d0, e, u0 = po.dg1d_initial(op, ninit=8, pert=1e-2)
    do, uo, to, ko, ts, dts = po.dg1d_run_rki(op, d0, e, u0, 10.0, max_steps=30)
    with fvwb.SolverDG1D(gpu_p(op)) as s:
        s.upload(d0, e, u0)
        s.advance(30)
        t, dtl, k = s.time()
        d, u = s.download()
    assert k == ko == 30 and t == to and dtl == dts[-1]
    same(d, do, "modes")
    same(u, uo, "u_init")


def test_run_in_pieces_equals_one_run():
    say("case: 1D DG RKi run in pieces")
    op = po.make_dg1d_params(64)
    d0, e, u0 = po.dg1d_initial(op, ninit=8, pert=1e-3)
    with fvwb.SolverDG1D(gpu_p(op)) as s:
        s.upload(d0, e, u0)
        s.advance(40)
        whole = s.download(), s.time()
        s.upload(d0, e, u0)
        for _ in range(4):
            s.advance(10)
        parts = s.download(), s.time()
    same(whole[0][0], parts[0][0], "modes")
    assert whole[1] == parts[1]


def test_run_error_paths():
    say("case: 1D DG run call order")
    with fvwb.SolverDG1D(fvwb.reference_params_dg1d(16)) as s:
        with pytest.raises(fvwb.FvwbError, match="upload"):
            s.advance(1)
        with pytest.raises(fvwb.FvwbError, match="upload"):
            s.download()
