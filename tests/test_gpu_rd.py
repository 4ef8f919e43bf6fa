"""The reference's 2D residual distribution (benchmark_rd.f90, solver 'RD1') on
the GPU, through the C ABI, against the oracle.  The device rounds statement by
statement like the oracle (no contraction, IEEE division and square root, the
reference's summation order at every node), so every comparison is bit for
bit."""
import sys

import numpy as np
import pytest

import fvwb
import pyoracle as po

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def params(nx, ny, bc=fvwb.RD_BC_TRANSMISSIVE, boxlen_y=1.0, cfl=0.5):
    return (fvwb.make_rd_params(nx, ny, bc=bc, boxlen_y=boxlen_y, cfl=cfl),
            po.make_rd_params(nx, ny, bc=bc, boxlen_y=boxlen_y, cfl=cfl))


def random_state(nx, ny, seed):
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.3 * rng.random((nx, ny))
    vx, vy = 0.4 * rng.standard_normal((2, nx, ny))
    p = 1.0 + 0.3 * rng.random((nx, ny))
    u = np.zeros((4, nx, ny), order="F")
    u[0], u[1], u[2] = rho, rho * vx, rho * vy
    u[3] = p / (po.GAMMA_REF - 1.0) + 0.5 * rho * (vx ** 2 + vy ** 2)
    return u


@pytest.mark.parametrize("bc", [fvwb.RD_BC_TRANSMISSIVE, fvwb.RD_BC_PERIODIC])
@pytest.mark.parametrize("shape", [(2, 2), (3, 5), (17, 9), (64, 48), (130, 70)])
def test_compute_update_matches_oracle(shape, bc):
    nx, ny = shape
    say(f"case: rd compute_update {nx}x{ny} bc={bc}")
    gp, op = params(nx, ny, bc, boxlen_y=0.7)
    u = random_state(nx, ny, nx * 31 + ny)
    diff = 1e-3 * np.asfortranarray(np.random.default_rng(ny).standard_normal((4, nx, ny)))
    with fvwb.SolverRD(gp) as s:
        for d in (np.zeros_like(diff), diff):
            got = s.compute_update(u, d, 2.5e-3)
            np.testing.assert_array_equal(got, po.rd_update(op, u, d, 2.5e-3))
        assert s.compute_max_speed(u) == po.rd_max_speed(op, u)


@pytest.mark.parametrize("ninit,bc", [(8, fvwb.RD_BC_TRANSMISSIVE), (1, fvwb.RD_BC_PERIODIC),
                                      (2, fvwb.RD_BC_TRANSMISSIVE)])
def test_run_matches_oracle(ninit, bc):
    nx, ny, nsteps = 49, 41, 25
    say(f"case: rd RD1 run ninit={ninit} bc={bc} {nx}x{ny} x {nsteps} steps")
    gp, op = params(nx, ny, bc, cfl=0.4)
    u0 = po.rd_initial(op, ninit, 1e-2)
    ref, t, n = po.rd_run(op, u0, np.inf, nsteps)
    with fvwb.SolverRD(gp) as s:
        s.upload(u0)
        s.advance(nsteps)
        got = s.download()
        tg, dtg, ng = s.time()
    assert ng == n == nsteps and tg == t
    np.testing.assert_array_equal(got, ref)


def test_t_end_clip_and_resume():
    """advance stops at t_end exactly, as evolve's dt = min(..., tend - t) does,
    and two advances equal one."""
    nx, ny = 33, 33
    say(f"case: rd t_end clip {nx}x{ny}")
    gp, op = params(nx, ny)
    u0 = po.rd_initial(op, 8, 1e-2)
    _, _, n1 = po.rd_run(op, u0, np.inf, 3)
    u = u0.copy(order="F")
    t = 0.0
    for _ in range(3):
        t += po.rd_step(op, u, t, np.inf)
    t_end = 0.5 * t
    ref, tr, nr = po.rd_run(op, u0, t_end, 100)
    with fvwb.SolverRD(gp) as s:
        s.upload(u0)
        s.advance(1, t_end)
        s.advance(100, t_end)
        tg, _, ng = s.time()
        got = s.download()
        s.advance(5, t_end)                     # a finished run does not move
        assert s.time()[2] == ng
    assert tg == tr == t_end and ng == nr
    np.testing.assert_array_equal(got, ref)


def test_large_grid_two_steps():
    n = 1025
    say(f"case: rd {n}x{n} two RD1 steps")
    gp, op = params(n, n)
    u0 = po.rd_initial(op, 8, 1e-3)
    ref, t, _ = po.rd_run(op, u0, np.inf, 2)
    with fvwb.SolverRD(gp) as s:
        s.upload(u0)
        s.advance(2)
        np.testing.assert_array_equal(s.download(), ref)
        assert s.time()[0] == t


def test_error_paths():
    say("case: rd error paths")
    for bad in (fvwb.make_rd_params(1, 4), fvwb.make_rd_params(4, 4, bc=1), fvwb.make_rd_params(4, 4, cfl=0.0)):
        with pytest.raises(fvwb.FvwbError) as e:
            fvwb.SolverRD(bad)
        assert e.value.code == fvwb.ERR_ARG
    gp, _ = params(8, 8)
    with fvwb.SolverRD(gp) as s:
        with pytest.raises(fvwb.FvwbError) as e:
            s.advance(1)
        assert e.value.code == fvwb.ERR_STATE
        with pytest.raises(fvwb.FvwbError):
            s.compute_update(np.zeros((4, 8, 7)), np.zeros((4, 8, 7)), 1e-3)
        u = random_state(8, 8, 1)
        u[:] = np.nan                                  # every speed NaN: no maximum
        s.upload(u)
        with pytest.raises(fvwb.FvwbError) as e:
            s.advance(1)
        assert e.value.code == fvwb.ERR_NONPHYSICAL
