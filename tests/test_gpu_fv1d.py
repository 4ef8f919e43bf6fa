"""1D HIP path (through the C ABI) against the oracle and the reference
fixtures.  The device paths round statement by statement like the oracle (no
contraction, IEEE division and square root), so the comparisons are bit for bit
except for compute_update_sr, whose pow comes from the device library (bounded
by a few ulps of the face fluxes times nx): the drop-in compute_update /
compute_max_speed of benchmark_1d.f90,
the configured residual, and whole runs of the LDS-resident batch kernel,
including the 1D gravity workload (256 cells, MUSCL-minmod, Rusanov, SSP-RK2,
walls, CFL 0.5, pulse 1e-3 at x=0.5 of width 0.05, t_end = 0.25)."""
import os
import sys

import numpy as np
import pytest

import fvwb
import pyoracle as po
from make_golden import bench1d_inputs

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def load(name):
    with np.load(os.path.join(GOLD, name + ".npz")) as z:
        return {k: z[k] for k in z.files}


def gpu_p(op):
    """fvwb.Params1D with the fields of an oracle Params1D."""
    return fvwb.Params1D(op.nvar, op.nx, op.bc, op.recon, op.integrator, op.nequilibrium, op.gamma, op.boxlen,
                         op.cfl, op.g)


def same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    if not np.array_equal(a, b):
        raise AssertionError(f"{what}: {np.count_nonzero(a != b)} entries differ, max {np.max(np.abs(a - b)):.3e}")


# ---- drop-in procedures --------------------------------------------------------------
def test_compute_update_matches_reference_golden_bitwise():
    say("case: 1D compute_update 128 cells, reference fixture")
    g = load("bench1d_update_128")
    p, u, w = bench1d_inputs()
    with fvwb.Solver1D(fvwb.reference_params_1d(128, 2)) as s:
        same(s.compute_update(u, w), g["dudt"], "dudt")
        same(s.compute_update(po.conservative_1d(p.gamma, w), w), g["dudt_eq"], "dudt at rest")
        assert s.compute_max_speed(u) == float(g["cmax"][0])


def test_polytropic_dropin_matches_reference_golden_bitwise():
    """compute_update and compute_max_speed against the reference built with
    nequilibrium 3 (the polytropic faces), bit for bit."""
    say("case: 1D compute_update nequilibrium 3, reference build")
    g = load("bench1d_update_neq3_128")
    p, u, w = bench1d_inputs(salt=3, nequilibrium=3)
    with fvwb.Solver1D(fvwb.reference_params_1d(128, 3)) as s:
        same(s.compute_update(u, w), g["dudt"], "dudt")
        same(s.compute_update(po.conservative_1d(p.gamma, w), w), g["dudt_eq"], "dudt at rest")
        assert s.compute_max_speed(u) == float(g["cmax"][0])


@pytest.mark.parametrize("nx,neq", [(37, 1), (256, 2), (1000, 3), (4, 2)])
def test_compute_update_matches_oracle_bitwise(nx, neq):
    say(f"case: 1D compute_update nx={nx} nequilibrium={neq}")
    op = po.make_params_1d(nx, nequilibrium=neq)
    u, w = po.perturbed_state_1d(op, amp_bump=1e-2, amp_noise=1e-3, salt=nx)
    u[1] += 0.05 * np.sin(np.arange(nx))
    with fvwb.Solver1D(fvwb.reference_params_1d(nx, neq)) as s:
        same(s.compute_update(u, w), po.ref1d_update(op, u, w), "dudt")
        assert s.compute_max_speed(u) == po.ref1d_max_speed(op, u)


# ---- compute_update_sr, the 'WB1' hydrostatic reconstruction ---------------------------
def sr_bound(u, gamma):
    """The device pow and the host's differ in the last place now and then;
    one ulp in a face state moves a flux by about ulp(|u| (|v|+c) + p), and the
    residual divides four of those by dx: the bound is 16 such ulps times nx."""
    w = np.stack([u[0], u[1] / u[0], (gamma - 1) * (u[2] - 0.5 * u[1] ** 2 / u[0])])
    c = np.sqrt(gamma * np.abs(w[2] / w[0]))
    f = np.max(np.sum(np.abs(u), axis=0) * (np.abs(w[1]) + c) + np.abs(w[2]))
    return 16 * np.finfo(float).eps * u.shape[1] * f


def close_sr(a, b, u, gamma, what):
    tol = sr_bound(u, gamma)
    err = np.max(np.abs(np.asarray(a) - np.asarray(b)))
    say(f"{what}: max|diff| {err:.3e} (bound {tol:.3e}), bitwise {np.mean(np.asarray(a) == b):.3f}")
    assert err <= tol, what


def test_compute_update_sr_matches_reference_golden():
    say("case: 1D compute_update_sr 128 cells, reference fixture")
    g = load("bench1d_update_sr_128")
    p = po.make_params_1d(128)
    with fvwb.Solver1D(fvwb.reference_params_1d(128, 2)) as s:
        for k in ("", "_eq", "_isen", "_jump"):
            close_sr(s.compute_update_sr(g["u" + k]), g["dudt" + k], g["u" + k], p.gamma, "dudt" + k)
        assert np.max(np.abs(s.compute_update_sr(g["u_isen"]))) <= 64 * np.finfo(float).eps * 128


@pytest.mark.parametrize("nx", [4, 37, 1000, 5000])
def test_compute_update_sr_matches_oracle(nx):
    say(f"case: 1D compute_update_sr nx={nx}")
    op = po.make_params_1d(nx)
    u, _ = po.perturbed_state_1d(op, amp_bump=1e-2, amp_noise=1e-3, salt=nx + 1)
    u[1] += 0.05 * np.sin(np.arange(nx))
    with fvwb.Solver1D(fvwb.reference_params_1d(nx, 2)) as s:
        close_sr(s.compute_update_sr(u), po.ref1d_update_sr(op, u), u, op.gamma, "dudt")


def test_compute_update_sr_keeps_isentropic_rest():
    """p = rho^gamma at rest under phi = x: dudt at rounding level on the device too."""
    nx = 2048
    op = po.make_params_1d(nx)
    x = (np.arange(nx) + 0.5) / nx
    w = np.zeros((3, nx), order="F")
    w[0] = (1 - ((op.gamma - 1) / op.gamma) * x) ** (1 / (op.gamma - 1))
    w[2] = w[0] ** op.gamma
    u = po.conservative_1d(op.gamma, w)
    with fvwb.Solver1D(fvwb.reference_params_1d(nx, 2)) as s:
        d = s.compute_update_sr(u)
    say(f"rest: max|dudt| {np.max(np.abs(d)):.3e}")
    assert np.max(np.abs(d)) <= 64 * np.finfo(float).eps * nx


# ---- the plain residuals: compute_update_fvm and fvm.f90's compute_update -------------
def test_plain_updates_match_reference_golden_bitwise():
    say("case: 1D compute_update_fvm 128 cells and fvm.f90 compute_update 200 cells, reference fixtures")
    g = load("bench1d_update_fvm_128")
    with fvwb.Solver1D(fvwb.reference_params_1d(128, 2)) as s:
        for k in ("", "_jump"):
            same(s.compute_update_fvm(g["u" + k]), g["dudt" + k], "fvm dudt" + k)
    g = load("fvm_update_200")
    q = fvwb.reference_params_1d(200, 2)
    q.bc = fvwb.BC1D_ZERO_GRADIENT
    with fvwb.Solver1D(q) as s:
        for k in ("_sod", "_smooth", "_poly"):
            same(s.fvm_compute_update(g["u" + k], 2), g["dudt" + k], "fvm.f90 dudt" + k)


@pytest.mark.parametrize("nx", [4, 37, 1000, 5000])
def test_plain_updates_match_oracle_bitwise(nx):
    say(f"case: plain 1D residuals nx={nx}")
    op = po.make_params_1d(nx)
    u, _ = po.perturbed_state_1d(op, amp_bump=1e-2, amp_noise=1e-3, salt=nx + 2)
    u[1] += 0.05 * np.sin(np.arange(nx))
    with fvwb.Solver1D(fvwb.reference_params_1d(nx, 2)) as s:
        same(s.compute_update_fvm(u), po.ref1d_update_fvm(op, u), "compute_update_fvm")
    for bc in (po.BC1D_PERIODIC, po.BC1D_ZERO_GRADIENT):
        q = fvwb.reference_params_1d(nx, 2)
        q.bc = bc
        with fvwb.Solver1D(q) as s:
            for src in (1, 2):
                same(s.fvm_compute_update(u, src), po.fvm_update(po.make_params_1d(nx, bc=bc), u, src),
                     f"fvm.f90 bc={bc} source={src}")


def test_plain_update_error_paths():
    say("case: fvm.f90 compute_update argument checks")
    u = po.conservative_1d(1.4, np.tile([[1.0], [0.0], [1.0]], (1, 16)))
    with fvwb.Solver1D(fvwb.reference_params_1d(16, 2)) as s:
        with pytest.raises(fvwb.FvwbError, match="source"):
            s.fvm_compute_update(u, 3)
    q = fvwb.reference_params_1d(16, 2)
    q.bc = po.BC1D_WALL
    with fvwb.Solver1D(q) as s:
        with pytest.raises(fvwb.FvwbError, match="bc 1"):
            s.fvm_compute_update(u, 2)


SCHEMES = [(r, bc, i) for r in (0, 1) for bc in (po.BC1D_ZERO_GRADIENT, po.BC1D_WALL) for i in (0, 1)]


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("neq,g", [(2, 1.0), (3, 0.8)])
def test_residual_matches_oracle_bitwise(scheme, neq, g):
    r, bc, integ = scheme
    say(f"case: 1D residual recon={r} bc={bc} neq={neq}")
    op = po.make_params_1d(150, recon=r, bc=bc, integrator=integ, nequilibrium=neq, g=g, gamma=1.4)
    u, w = po.perturbed_state_1d(op, amp_bump=1e-2, x0=0.1, amp_noise=1e-3, salt=3)
    with fvwb.Solver1D(gpu_p(op)) as s:
        same(s.compute_residual(u, w), po.fv1d_residual(op, u, w), "residual")


# ---- device-resident runs --------------------------------------------------------------
def test_gravity_workload_run_matches_oracle_bitwise():
    """The 1D gravity workload to t = 0.25, whole run in one launch."""
    say("case: 1D workload 256 cells to t=0.25")
    op = po.make_params_1d(256, recon=po.RECON_MUSCL_MINMOD, bc=po.BC1D_WALL, integrator=po.RK2_HEUN,
                           nequilibrium=2, gamma=1.4, cfl=0.5, g=1.0)
    u, w = po.perturbed_state_1d(op, amp_bump=1e-3, x0=0.5, width=0.05)
    uo, to, no, dto = po.fv1d_advance(op, u, w, t_end=0.25)
    with fvwb.Solver1D(gpu_p(op)) as s:
        s.upload(u, w)
        s.advance(t_end=0.25)
        t, dt, n = s.time()
        assert n == no and t == to == 0.25 and dt == dto
        same(s.download(), uo, "state")


@pytest.mark.parametrize("scheme", SCHEMES)
def test_every_scheme_run_matches_oracle_bitwise(scheme):
    r, bc, integ = scheme
    say(f"case: 1D run recon={r} bc={bc} integ={integ}")
    op = po.make_params_1d(300, recon=r, bc=bc, integrator=integ, nequilibrium=3, g=0.9, gamma=1.4, cfl=0.45)
    u, w = po.perturbed_state_1d(op, amp_bump=5e-2, x0=0.3, width=0.04, amp_noise=1e-3)
    uo, to, no, _ = po.fv1d_advance(op, u, w, t_end=0.1)
    with fvwb.Solver1D(gpu_p(op)) as s:
        s.upload(u, w)
        s.advance(t_end=0.1)
        assert s.time()[2] == no
        same(s.download(), uo, "state")


def test_batch_of_uneven_instances_matches_oracle_bitwise():
    """48 instances of 4..2048 cells, each with its own g, amplitude and
    clock, one workgroup each; every instance equals its own oracle run."""
    say("case: 1D batch of 48 uneven instances")
    rng = np.random.default_rng(11)
    nxs = np.concatenate([[4, 5, 255, 256, 257, 1024, 1025, 2048], rng.integers(4, 2049, 40)])
    ops, states = [], []
    for k, nx in enumerate(nxs):
        op = po.make_params_1d(int(nx), recon=po.RECON_MUSCL_MINMOD, bc=po.BC1D_WALL, integrator=po.SSPRK3,
                               nequilibrium=2, g=float(rng.uniform(0.5, 2.0)), gamma=1.4, cfl=0.5)
        u, w = po.perturbed_state_1d(op, amp_bump=float(rng.uniform(1e-4, 1e-2)), x0=float(rng.uniform(0.2, 0.8)),
                                     width=0.05, amp_noise=1e-4, salt=k)
        ops.append(op)
        states.append((u, w))
    with fvwb.Solver1D([gpu_p(op) for op in ops]) as s:
        for k, (u, w) in enumerate(states):
            s.upload(u, w, inst=k)
        s.advance(nsteps=30)
        for k, op in enumerate(ops):
            u, w = states[k]
            uo, to, no, dto = po.fv1d_advance(op, u, w, max_steps=30)
            t, dt, n = s.time(k)
            assert n == no == 30 and t == to and dt == dto, k
            same(s.download(k), uo, f"instance {k} (nx={op.nx})")


def test_resting_atmospheres_stay_exactly_at_rest():
    say("case: 1D equilibria at rest")
    ops = [po.make_params_1d(nx, recon=1, bc=po.BC1D_WALL, integrator=po.SSPRK3, nequilibrium=neq, g=g, gamma=1.4)
           for nx, neq, g in ((64, 2, 1.0), (500, 3, 1.0), (2048, 2, 3.0))]
    with fvwb.Solver1D([gpu_p(op) for op in ops]) as s:
        s.init_gaussian(np.zeros(len(ops)))
        before = [s.download(k) for k in range(len(ops))]
        s.advance(nsteps=200)
        for k in range(len(ops)):
            assert s.time(k)[2] == 200
            same(s.download(k), before[k], "rest")


def test_device_initial_condition_matches_oracle_set_up():
    say("case: 1D init_gaussian")
    op = po.make_params_1d(256, recon=1, bc=po.BC1D_WALL, gamma=1.4)
    u, w = po.perturbed_state_1d(op, amp_bump=1e-3, x0=0.5, width=0.05)
    with fvwb.Solver1D(gpu_p(op)) as s:
        s.init_gaussian(1e-3, 0.5, 0.05)
        np.testing.assert_allclose(s.download(), u, rtol=4e-16, atol=0)


def test_error_paths():
    say("case: 1D error paths")
    ok = po.make_params_1d(64, recon=1, bc=po.BC1D_WALL, gamma=1.4)
    with fvwb.Solver1D(gpu_p(ok)) as s:
        with pytest.raises(fvwb.FvwbError) as e:
            s.advance(5)                      # no state yet
        assert e.value.code == fvwb.ERR_STATE
        u, w = po.perturbed_state_1d(ok, amp_bump=-50.0, x0=0.5, width=0.05)   # negative pressure
        s.upload(u, w)
        with pytest.raises(fvwb.FvwbError) as e:
            s.advance(5)
        assert e.value.code == fvwb.ERR_NONPHYSICAL
        assert s.time()[2] == 0
    big = po.make_params_1d(2049, recon=1, bc=po.BC1D_WALL, gamma=1.4)
    with fvwb.Solver1D(gpu_p(big)) as s:
        u, w = po.perturbed_state_1d(big)
        s.upload(u, w)
        with pytest.raises(fvwb.FvwbError) as e:
            s.advance(1)
        assert e.value.code == fvwb.ERR_ARG
        # the per-stage drop-in has no size limit
        ref = fvwb.reference_params_1d(2049, 2)
    with fvwb.Solver1D(ref) as s:
        same(s.compute_update(u, w), po.ref1d_update(po.make_params_1d(2049), u, w), "dudt 2049")
    bad = po.make_params_1d(64, bc=po.BC1D_PERIODIC, gamma=1.4)
    with fvwb.Solver1D(gpu_p(bad)) as s:
        u, w = po.perturbed_state_1d(bad)
        s.upload(u, w)
        with pytest.raises(fvwb.FvwbError):
            s.advance(1)
