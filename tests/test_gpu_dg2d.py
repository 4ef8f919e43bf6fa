"""P2 modal DG on the GPU (FP32, projections on the matrix cores) against the
FP64 oracle: residuals, the TVB limiter, whole SSP-RK3 runs, the Fourier
initial state, exact rest and the clock.

Tolerances: the device works in FP32 on the deviation (about 6e-8 relative to
the deviation, never to the O(1) state), so residuals agree to 2e-5 of their
maximum and states after tens of steps to 1e-5 of the deviation's maximum; the
resting atmosphere stays exactly zero."""
import sys

import numpy as np
import pytest

import fvwb
import pyoracle as po

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_p(op):
    return fvwb.DGParams(op.nvar, op.nx, op.ny, op.degree, op.flux, op.limiter, op.bc, op.integrator, op.gamma,
                         op.boxlen_x, op.boxlen_y, op.cfl, op.rho0, op.p0, op.g, op.tvb_m)


def smooth_noisy(op, amp=1e-3, noise=1e-4, seed=1):
    d = po.dg2d_fourier_state(op, amp=amp)
    rng = np.random.default_rng(seed)
    d = d + noise * rng.normal(size=d.shape) * np.array([1.0, 0.5, 0.5, 2.0])[:, None, None, None]
    return np.asfortranarray(d)


@pytest.mark.parametrize("shape", [(16, 16), (40, 37), (17, 5), (64, 64), (2, 33)])
@pytest.mark.parametrize("flux", (po.FLUX_HLL, po.FLUX_RUSANOV, po.FLUX_HLLC))
def test_residual_matches_oracle(shape, flux):
    nx, ny = shape
    say(f"case: DG residual {nx}x{ny} flux={flux}")
    op = po.make_dg_params(nx, ny, flux=flux, g=1.3, rho0=1.1, p0=0.9)
    d = smooth_noisy(op, seed=nx + ny)
    ref = po.dg2d_residual(op, d)
    with fvwb.SolverDG(gpu_p(op)) as s:
        got = s.compute_residual(d)
    err = np.max(np.abs(got - ref))
    say(f"   max |L| {np.max(np.abs(ref)):.3e}, max err {err:.3e}")
    assert err <= 2e-5 * np.max(np.abs(ref))


@pytest.mark.parametrize("shape", [(1000, 1030), (2000, 301)])
def test_residual_matches_oracle_large(shape):
    """Grids large enough that the sweep runs segments of many rows (25 and 15
    rows on a 256-CU device at 3 waves per SIMD) with short last segments, and
    strips whose last lanes fall past nx."""
    nx, ny = shape
    say(f"case: DG residual {nx}x{ny} (multi-row segments)")
    op = po.make_dg_params(nx, ny, g=1.3, rho0=1.1, p0=0.9)
    d = np.asfortranarray(po.dg2d_fourier_state(op, amp=1e-3))
    ref = po.dg2d_residual(op, d)
    with fvwb.SolverDG(gpu_p(op)) as s:
        got = s.compute_residual(d)
    e = np.abs(got - ref)
    err = float(e.max())
    where = np.unravel_index(np.argmax(e), e.shape)
    row_err = e.max(axis=(0, 1, 3))
    say(f"   max |L| {np.max(np.abs(ref)):.3e}, max err {err:.3e} at (var, i, j, mode) {where}, "
        f"median row max {np.median(row_err):.3e}")
    # FP32 flux rounding (about 2^-24 of the deviation flux) is divided by dx and
    # dy, so the bound grows with the resolution; a wrong trace or face anywhere
    # (a segment or strip edge) gives an error of the order of |L| itself
    dmax = float(np.max(np.abs(d)))
    assert err <= 2e-5 * np.max(np.abs(ref)) + 2e-6 * dmax * (nx / op.boxlen_x + ny / op.boxlen_y)
    # no row stands out: segment boundaries are as accurate as interior rows
    assert row_err.max() <= 50 * np.median(row_err)


def test_run_matches_oracle_multirow_segments():
    say("case: DG 3 SSP-RK3 steps on 1000x400 (10-row segments)")
    op = po.make_dg_params(1000, 400, limiter=po.DG_LIMITER_TVB, g=1.0)
    d = po.dg2d_limit(op, po.dg2d_fourier_state(op, amp=1e-4))
    ref, t, n, dt = po.dg2d_advance(op, d, max_steps=3)
    with fvwb.SolverDG(gpu_p(op)) as s:
        s.upload(d)
        s.advance(3)
        tg, dtg, ng = s.time()
        got = s.download()
    assert ng == n == 3 and abs(tg - t) <= 1e-6 * t
    scale = np.max(np.abs(ref))
    err = np.max(np.abs(got - ref))
    say(f"   max |d| {scale:.3e}, max err {err:.3e}")
    assert err <= 1e-5 * scale


@pytest.mark.parametrize("integrator,flux", [(po.DG_SSPRK3, po.FLUX_HLL), (po.DG_SSPRK54, po.FLUX_HLL),
                                             (po.DG_SSPRK3, po.FLUX_HLLC)])
def test_rest_is_exact(integrator, flux):
    say(f"case: DG rest integrator={integrator} flux={flux}")
    op = po.make_dg_params(37, 29, g=2.0, integrator=integrator, flux=flux)
    z = np.zeros((4, 37, 29, 6), order="F")
    with fvwb.SolverDG(gpu_p(op)) as s:
        assert np.count_nonzero(s.compute_residual(z)) == 0
        s.upload(z)
        s.advance(25)
        assert s.time()[2] == 25
        assert np.count_nonzero(s.download()) == 0


def test_limiter_matches_oracle():
    say("case: DG TVB limiter on a step")
    op = po.make_dg_params(32, 24, tvb_m=50.0)
    X, Y = po.dg2d_points(op)
    nodal = np.zeros((4, 32, 24, 9), order="F")
    nodal[3] = np.where(X < 0.37, 0.5, 0.0)     # slopes well above M dx^2 = 0.049
    nodal[0] = np.where(X + Y < 0.8, 0.05, 0.0)
    d = po.dg2d_project(op, nodal)
    ref = po.dg2d_limit(op, d)
    assert np.any(ref != d)
    with fvwb.SolverDG(gpu_p(op)) as s:
        s.upload(d)
        got = s.download()
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-7 * np.max(np.abs(d)))


def gauss_density_norms(op, a, b):
    """Density error of two mode sets at the 3x3 Gauss points of every cell:
    the L1 norm weighted as compute_error weights it (2d/benchmark_2d_dg.f90:
    45-59: sum over cells and points of |d| w_a w_b dx dy / 4) and the maximum."""
    diff = np.abs(po.dg2d_nodal(op, a)[0] - po.dg2d_nodal(op, b)[0])
    _, wq = po.dg_gl_quadrature(3)
    wts = np.outer(wq, wq).ravel(order="F")
    dxdy = (op.boxlen_x / op.nx) * (op.boxlen_y / op.ny)
    return float(np.sum(diff * wts[None, None, :]) * 0.25 * dxdy), float(diff.max())


@pytest.mark.parametrize("integrator", (po.DG_SSPRK3, po.DG_SSPRK54))
@pytest.mark.parametrize("limiter", (po.DG_LIMITER_NONE, po.DG_LIMITER_TVB, po.DG_LIMITER_CHARACTERISTIC))
def test_run_matches_oracle(limiter, integrator):
    say(f"case: DG 30 steps limiter={limiter} integrator={integrator}")
    op = po.make_dg_params(48, 40, limiter=limiter, g=1.0, integrator=integrator)
    d = po.dg2d_limit(op, po.dg2d_fourier_state(op, amp=1e-4))
    ref, t, n, dt = po.dg2d_advance(op, d, max_steps=30)
    with fvwb.SolverDG(gpu_p(op)) as s:
        s.upload(d)
        s.advance(30)
        tg, dtg, ng = s.time()
        got = s.download()
    assert ng == n == 30 and abs(tg - t) <= 1e-6 * t
    scale = np.max(np.abs(ref))
    err = np.max(np.abs(got - ref))
    l1 = np.mean(np.abs(got[0, :, :, 0] - ref[0, :, :, 0]))
    l1g, linfg = gauss_density_norms(op, got, ref)
    say(f"   max |d| {scale:.3e}, max err {err:.3e}, L1 density (means) {l1:.3e}, "
        f"Gauss-weighted L1 {l1g:.3e}, Linf at the Gauss points {linfg:.3e}")
    assert err <= 1e-5 * scale and l1 <= 1e-6 and l1g <= 1e-6 and linfg <= 1e-5 * scale


@pytest.mark.parametrize("shape", [(70, 45), (300, 130)])
def test_tvb_positivity_equals_tvb_away_from_vacuum(shape):
    """The TVB-plus-positivity limiter pass is a separate compiled variant of
    the TVB pass.  On a smooth state far from vacuum its positivity part changes
    nothing, so whole runs with either limiter must agree bit for bit.  M = 0 and
    noise make the TVB part act on many cells, across strip and segment edges."""
    nx, ny = shape
    say(f"case: DG TVB vs TVB+positivity {nx}x{ny}")
    op = po.make_dg_params(nx, ny, tvb_m=0.0, g=1.0)
    d = smooth_noisy(op, amp=1e-3, noise=1e-4, seed=nx)
    lim = po.dg2d_limit(op, d)
    assert np.count_nonzero(np.any(lim != d, axis=(0, 3))) > nx * ny // 10
    out = {}
    for limiter in (po.DG_LIMITER_TVB, po.DG_LIMITER_TVB_POSITIVITY):
        op.limiter = limiter
        with fvwb.SolverDG(gpu_p(op)) as s:
            s.upload(lim)
            s.advance(12)
            out[limiter] = (s.download(), s.time())
    a, ta = out[po.DG_LIMITER_TVB]
    b, tb = out[po.DG_LIMITER_TVB_POSITIVITY]
    say(f"   max |a - b| {np.max(np.abs(a - b)):.3e}, t {ta[0]:.6e} / {tb[0]:.6e}")
    assert ta == tb
    np.testing.assert_array_equal(a, b)


def test_fourier_initial_state_and_t_end():
    say("case: DG init_fourier, t_end clip")
    op = po.make_dg_params(32, 32)
    a, mk, nk, th = fvwb.dg_fourier_terms()
    for x, y in zip((a, mk, nk, th), po.fourier_modes()):
        np.testing.assert_array_equal(x, y)
    ref = po.dg2d_limit(op, po.dg2d_fourier_state(op, amp=1e-4))
    t_end = 3e-3
    with fvwb.SolverDG(gpu_p(op)) as s:
        s.init_fourier(a, mk.astype(float), nk.astype(float), th, 1e-4)
        np.testing.assert_allclose(s.download(), ref, rtol=0, atol=2e-7 * np.max(np.abs(ref)))
        s.advance(1000, t_end)
        t, dt, n = s.time()
        s.advance(5, t_end)                 # a finished instance does not move
        assert s.time() == (t, dt, n)
    _, to, no, _ = po.dg2d_advance(op, ref, t_end=t_end)
    assert t == t_end and n == no


@pytest.mark.parametrize("integrator", (po.DG_SSPRK3, po.DG_SSPRK54))
def test_hllc_run_matches_oracle(integrator):
    say(f"case: DG HLLC 30 steps with TVB, integrator={integrator}")
    op = po.make_dg_params(48, 40, flux=po.FLUX_HLLC, g=1.0, integrator=integrator)
    d = po.dg2d_limit(op, po.dg2d_fourier_state(op, amp=1e-4))
    ref, t, n, _ = po.dg2d_advance(op, d, max_steps=30)
    with fvwb.SolverDG(gpu_p(op)) as s:
        s.upload(d)
        s.advance(30)
        tg, _, ng = s.time()
        got = s.download()
    assert ng == n == 30 and abs(tg - t) <= 1e-6 * t
    scale = np.max(np.abs(ref))
    err = np.max(np.abs(got - ref))
    say(f"   max |d| {scale:.3e}, max err {err:.3e}")
    assert err <= 1e-5 * scale


def test_hllc_contact_and_limiter_match_oracle():
    """A density and shear contact plus a pressure step: every HLLC region and the
    limiter are exercised."""
    say("case: DG HLLC across a contact and a step")
    op = po.make_dg_params(32, 24, flux=po.FLUX_HLLC, tvb_m=50.0)
    X, Y = po.dg2d_points(op)
    nodal = np.zeros((4, 32, 24, 9), order="F")
    nodal[0] = np.where(Y < 0.45, 0.2, -0.1)                 # density jump
    nodal[1] = np.where(Y < 0.45, 0.05, -0.05)               # shear
    nodal[3] = 0.05 * np.where(X < 0.37, 1.0, 0.0)
    d = po.dg2d_limit(op, po.dg2d_project(op, nodal))
    ref = po.dg2d_residual(op, d)
    with fvwb.SolverDG(gpu_p(op)) as s:
        got = s.compute_residual(d)
        r, t, n, _ = po.dg2d_advance(op, d, max_steps=5)
        s.upload(d)
        s.advance(5)
        g5 = s.download()
    err = np.max(np.abs(got - ref))
    say(f"   residual: max |L| {np.max(np.abs(ref)):.3e}, max err {err:.3e}")
    assert err <= 2e-5 * np.max(np.abs(ref))
    err5 = np.max(np.abs(g5 - r))
    say(f"   5 steps: max |d| {np.max(np.abs(r)):.3e}, max err {err5:.3e}")
    assert err5 <= 1e-5 * np.max(np.abs(r))


def test_ssp54_t_end_and_limited_steps():
    say("case: DG SSP(5,4) with the limiter active, to t_end")
    op = po.make_dg_params(32, 24, tvb_m=50.0, integrator=po.DG_SSPRK54)
    X, Y = po.dg2d_points(op)
    nodal = np.zeros((4, 32, 24, 9), order="F")
    nodal[3] = 0.05 * np.where(X < 0.37, 1.0, 0.0)      # a pressure step the limiter acts on
    d = po.dg2d_limit(op, po.dg2d_project(op, nodal))
    t_end = 4e-3
    ref, t, n, _ = po.dg2d_advance(op, d, t_end=t_end)
    with fvwb.SolverDG(gpu_p(op)) as s:
        s.upload(d)
        s.advance(1000, t_end)
        tg, _, ng = s.time()
        got = s.download()
    assert tg == t_end and ng == n
    scale = np.max(np.abs(ref))
    err = np.max(np.abs(got - ref))
    say(f"   {n} steps, max |d| {scale:.3e}, max err {err:.3e}")
    assert err <= 1e-5 * scale


def test_error_paths():
    say("case: DG error paths")
    with pytest.raises(fvwb.FvwbError):
        fvwb.SolverDG(gpu_p(po.make_dg_params(8, 8, degree=1)))
    with pytest.raises(fvwb.FvwbError):
        fvwb.SolverDG(gpu_p(po.make_dg_params(8, 8, integrator=2)))
    op = po.make_dg_params(8, 8)
    with fvwb.SolverDG(gpu_p(op)) as s:
        with pytest.raises(fvwb.FvwbError) as e:
            s.advance(1)
        assert e.value.code == fvwb.ERR_STATE
        d = np.zeros((4, 8, 8, 6), order="F")
        d[3, :, :, 0] = -10.0                 # negative pressure
        with pytest.raises(fvwb.FvwbError) as e:
            s.upload(d)
        assert e.value.code == fvwb.ERR_NONPHYSICAL


# ---- positivity limiter (compute_positivity, 2d/limiters.f90:478-654) -------------------
def equilibrium_modes(op):
    X, Y = po.dg2d_points(op)
    e = np.exp(-(op.rho0 * op.g / op.p0) * (X + Y))
    nodal = np.zeros((4,) + e.shape, order="F")
    nodal[0] = op.rho0 * e
    nodal[3] = op.p0 * e / (op.gamma - 1)
    return po.dg2d_project(op, nodal)


def outflow_state(op, speed=2.5):
    """A deviation that empties the box's centre: the equilibrium density with a
    radial velocity of the given size, projected on P2 (the momentum and the
    kinetic energy are the deviation)."""
    X, Y = po.dg2d_points(op)
    k = op.rho0 * op.g / op.p0
    r = op.rho0 * np.exp(-k * (X + Y))
    dx, dy = X - 0.5, Y - 0.5
    rr = np.sqrt(dx * dx + dy * dy) + 1e-12
    prof = speed * np.tanh(rr / 0.05) * np.exp(-(rr / 0.3) ** 2)
    vx, vy = prof * dx / rr, prof * dy / rr
    nodal = np.zeros((4,) + X.shape, order="F")
    nodal[1] = r * vx
    nodal[2] = r * vy
    nodal[3] = 0.5 * r * (vx * vx + vy * vy)
    return po.dg2d_project(op, nodal)


@pytest.mark.parametrize("limiter", (po.DG_LIMITER_POSITIVITY, po.DG_LIMITER_TVB_POSITIVITY))
def test_positivity_limiter_matches_oracle(limiter):
    """One pass of the limiter (on upload) over a deviation with vacuum-like
    points: the device's FP32 limiter equals the oracle's FP64 one on the
    FP32-rounded input, to FP32 rounding of the full state."""
    say(f"case: DG positivity limiter={limiter}")
    op = po.make_dg_params(24, 20, limiter=limiter, tvb_m=0.0)
    E = equilibrium_modes(op)
    rng = np.random.default_rng(478)
    d = np.asfortranarray(rng.normal(size=(4, 24, 20, 6)) * 0.5 * np.abs(E[..., :1]))
    d[..., 0] = 0.0
    d = np.asfortranarray(d.astype(np.float32).astype(np.float64))
    ref = po.dg2d_limit(op, d)
    assert np.any(ref != d)
    with fvwb.SolverDG(gpu_p(op)) as s:
        s.upload(d)
        got = s.download()
    err = np.max(np.abs(got - ref))
    say(f"   limited cells {np.count_nonzero(np.any(ref != d, axis=(0, 3)))}, max err {err:.3e}")
    np.testing.assert_array_equal(got[..., 0], d[..., 0])
    assert err <= 2e-5 * np.max(np.abs(E))


def test_positivity_limiter_keeps_a_near_vacuum_run_alive():
    """A strong outflow from the centre (speed 6, five times the sound speed)
    empties it: without a limiter, and with TVB alone, the P2 run stops with a
    non-positive state; with the positivity limiter it runs all 300 steps with
    positive density and pressure at every volume point."""
    steps = 300

    def run(limiter):
        op = po.make_dg_params(48, 48, limiter=limiter, cfl=0.5)
        with fvwb.SolverDG(gpu_p(op)) as s:
            s.upload(outflow_state(op, speed=6.0))
            for _ in range(steps // 50):
                try:
                    s.advance(50)
                except fvwb.FvwbError as e:
                    assert "rho<=0" in str(e)
                    return op, s.time()[2], None
            return op, s.time()[2], s.download()

    _, n0, got0 = run(po.DG_LIMITER_NONE)
    _, n1, got1 = run(po.DG_LIMITER_TVB)
    op, n, got = run(po.DG_LIMITER_POSITIVITY)
    assert got is not None and n == steps
    nodal = po.dg2d_nodal(op, got)
    X, Y = po.dg2d_points(op)
    e = np.exp(-(X + Y))
    rho = e + nodal[0]
    pr = (op.gamma - 1) * (e / (op.gamma - 1) + nodal[3] - 0.5 * (nodal[1] ** 2 + nodal[2] ** 2) / rho)
    say(f"   unlimited stopped by step {n0}, TVB by {n1}; positivity: {n} steps, min rho {rho.min():.3e}, "
        f"min p {pr.min():.3e}")
    assert got0 is None and got1 is None
    assert rho.min() > 0 and pr.min() > 0



# ---- characteristic limiting --------------------------------------------------------------
@pytest.mark.parametrize("tvb_m", [0.0, 50.0])
def test_characteristic_limiter_matches_oracle(tvb_m):
    """One pass (on upload) over steps in density, pressure and velocity: the
    device's FP32 characteristic limiter equals the oracle's on the
    FP32-rounded input to FP32 rounding, cell by cell (the same cells limited)."""
    say(f"case: DG characteristic limiter M={tvb_m}")
    op = po.make_dg_params(32, 24, tvb_m=tvb_m, limiter=po.DG_LIMITER_CHARACTERISTIC)
    X, Y = po.dg2d_points(op)
    nodal = np.zeros((4, 32, 24, 9), order="F")
    nodal[3] = np.where(X < 0.37, 0.5, 0.0)
    nodal[0] = np.where(X + Y < 0.8, 0.05, 0.0)
    nodal[1] = 0.2 * np.where(Y > 0.5, 1.0, -1.0)
    d = np.asfortranarray(po.dg2d_project(op, nodal).astype(np.float32).astype(np.float64))
    ref = po.dg2d_limit(op, d)
    with fvwb.SolverDG(gpu_p(op)) as s:
        s.upload(d)
        got = s.download()
    lim_ref = np.any(ref != d, axis=(0, 3))
    lim_got = np.any(got != d, axis=(0, 3))
    err = np.max(np.abs(got - ref))
    say(f"   limited cells {lim_ref.sum()} (device {lim_got.sum()}), max err {err:.3e}")
    assert lim_ref.sum() > 0 and np.array_equal(lim_ref, lim_got)
    assert err <= 2e-6 * np.max(np.abs(d))


def test_characteristic_limiter_keeps_rest_exact():
    say("case: DG rest with the characteristic limiter")
    op = po.make_dg_params(24, 24, limiter=po.DG_LIMITER_CHARACTERISTIC_POSITIVITY, integrator=po.DG_SSPRK54)
    z = np.zeros((4, 24, 24, 6), order="F")
    with fvwb.SolverDG(gpu_p(op)) as s:
        s.upload(z)
        s.advance(20)
        assert np.count_nonzero(s.download()) == 0
