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
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    if not np.array_equal(a, b):
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
