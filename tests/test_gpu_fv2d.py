"""HIP path (through the C ABI) against the CPU oracle and the reference
fixtures: the drop-in compute_update_exact / compute_max_speed / evolve on the
reference's own 300x300 set-up, every scheme combination, the FP64 plain and
FP32 deviation stepping, exact well-balancing, edge shapes and error paths.

Tolerances: FP64 results differ from the oracle by FMA contraction, the
separable equilibrium tables (e^{-kx} e^{-ky} instead of e^{-k(x+y)}) and the
deviation storage, all at the level of a few ulp of the O(1) face fluxes
divided by dx: |d dudt| <= 64*eps*2/dx.  FP32: L1 density error <= 1e-6 (the
BASELINE accuracy target) and max error <= 1e-6."""
import itertools
import math
import os
import sys

import numpy as np
import pytest

import fvwb
import pyoracle as po
from make_golden import bench2d_inputs, summarize

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SCHEMES = list(itertools.product((0, 1), (0, 1), (0, 1)))   # recon, flux, bc


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def load(name):
    with np.load(os.path.join(GOLD, name + ".npz")) as z:
        return {k: z[k] for k in z.files}


def gpu_params(nx, ny, recon=0, flux=0, bc=0, integ=0, prec=0, **kw):
    return fvwb.make_params(nx, ny, recon=recon, flux=flux, bc=bc, integrator=integ, precision=prec, **kw)


def orc_params(gp):
    return po.make_params(gp.nx, gp.ny, bc=gp.bc, recon=gp.recon, flux=gp.flux,
                          integrator=gp.integrator, gamma=gp.gamma, boxlen_x=gp.boxlen_x,
                          boxlen_y=gp.boxlen_y, cfl=gp.cfl, rho0=gp.rho0, p0=gp.p0, g=gp.g)


def res_tol(nx):
    return 64 * EPS * 2.0 * nx


# ---- the reference's own set-up (benchmark_2d.f90, 300x300) ----------------------
def test_compute_update_exact_matches_reference_golden_and_oracle():
    say("case: compute_update_exact 300x300 reference set-up")
    g = load("bench2d_update_exact_300")
    p, u, w = bench2d_inputs(salt=int(g["salt"]))
    with fvwb.Solver2D(fvwb.reference_params(300, 300, 2)) as s:
        d = s.compute_update_exact(u, w)
    orc = po.ref2d_update_exact(p, 2, u, w)
    assert np.max(np.abs(d - orc)) <= res_tol(300)
    sm = summarize(d)
    assert np.max(np.abs(sm["sample"] - g["sample"])) <= res_tol(300)
    # the frozen ring is exactly zero (benchmark_2d.f90:611-614)
    for a in (d[:, 0, :], d[:, -1, :], d[:, :, 0], d[:, :, -1]):
        assert np.count_nonzero(a) == 0


def test_compute_update_plain_matches_reference_golden_bitwise():
    """compute_update (benchmark_2d.f90:370-463), the plain non-WB residual: the
    device result equals the reference's (golden summary) and the oracle's bit for
    bit on the reference's 300x300 set-up."""
    say("case: compute_update (plain) 300x300 reference set-up")
    g = load("bench2d_update_plain_300")
    p, u, w = bench2d_inputs(salt=int(g["salt"]), amp_bump=float(g["amp_bump"]), amp_noise=float(g["amp_noise"]))
    with fvwb.Solver2D(fvwb.reference_params(300, 300, 2)) as s:
        d = s.compute_update(u, w)
    np.testing.assert_array_equal(d, po.ref2d_update_plain(p, u))
    sm = summarize(d)
    for k in ("sample", "sum", "abssum", "maxabs"):
        np.testing.assert_array_equal(sm[k], g[k])


@pytest.mark.parametrize("shape", [(64, 48), (17, 90), (4, 5)])
def test_compute_update_plain_other_shapes_bitwise(shape):
    nx, ny = shape
    say(f"case: compute_update (plain) {nx}x{ny}")
    gp = fvwb.reference_params(nx, ny, 2)
    p = orc_params(gp)
    rng = np.random.default_rng(nx * 100 + ny)
    w = po.equilibrium_primitive(p)
    u = po.conservative(p.gamma, w)
    u = u * (1 + 1e-2 * rng.normal(size=u.shape))
    with fvwb.Solver2D(gp) as s:
        d = s.compute_update(u)
    np.testing.assert_array_equal(d, po.ref2d_update_plain(p, u))


def test_compute_update_exact_equilibrium_is_bitwise_zero():
    say("case: compute_update_exact at equilibrium 300x300")
    p, u, w = bench2d_inputs()
    ueq = po.conservative(p.gamma, w)
    with fvwb.Solver2D(fvwb.reference_params(300, 300, 2)) as s:
        d = s.compute_update_exact(ueq, w)
    assert np.count_nonzero(d) == 0


@pytest.mark.parametrize("shape", [(64, 48), (40, 90)])
def test_compute_update_exact_non_square_keeps_reference_face_positions(shape):
    """On nx != ny the reference places the y faces at (j-1)*dx
    (benchmark_2d.f90:507-514); the drop-in follows it, while the configured
    residual uses (j-1)*dy, so the two differ."""
    nx, ny = shape
    say(f"case: compute_update_exact {nx}x{ny} reference set-up")
    p = po.make_params(nx, ny)
    u, w = po.perturbed_state(p, amp_bump=1e-3, amp_noise=1e-4, salt=4)
    with fvwb.Solver2D(fvwb.reference_params(nx, ny, 2)) as s:
        d = s.compute_update_exact(u, w)
        r = s.compute_residual(u, w)
    assert np.max(np.abs(d - po.ref2d_update_exact(p, 2, u, w))) <= res_tol(max(nx, ny))
    assert np.max(np.abs(r - po.fv2d_residual(p, u, w))) <= res_tol(max(nx, ny))
    assert np.max(np.abs(d - r)) > 1e3 * res_tol(max(nx, ny))


def test_compute_max_speed_matches_reference_golden():
    say("case: compute_max_speed 300x300")
    g = load("bench2d_max_speed_300")
    p, u, w = bench2d_inputs(salt=int(g["salt"]))
    with fvwb.Solver2D(fvwb.reference_params(300, 300, 2)) as s:
        c = s.compute_max_speed(u)
    assert abs(c - float(g["cmax"][0])) <= 4 * EPS


def test_evolve_matches_reference_golden():
    """The golden's w_eq is the closed form rounded through FP32.  The device
    path takes the wave speed for dt from the closed-form centre equilibrium
    plus the stored deviation, so dt differs from the reference's (full u) by
    ~FP32 rounding of w_eq (~1e-8 relative); over three steps the state then
    differs by ~1e-11."""
    say("case: evolve (Heun) 300x300, three steps, golden input")
    g = load("bench2d_evolve_300")
    p, u, w = bench2d_inputs(salt=int(g["salt"]))
    ug, t, n = fvwb.evolve(u, w, fvwb.reference_params(300, 300, 2), t_end=float(g["tend"]))
    assert n == 3
    sm = summarize(ug - po.conservative(p.gamma, w))
    assert np.max(np.abs(sm["sample"] - g["sample"])) <= 1e-10


def test_dropins_with_dx_ne_dy_match_reference_golden():
    """The reference compiled at 300 x 200 (dx != dy): the drop-ins
    compute_update_exact (y faces at (j-1)*dx, benchmark_2d.f90:513) and
    compute_max_speed.  The resident loop integrates the residual with the y
    faces at (j-1)*dy instead (fvwb_compute_residual): it equals the oracle's
    run of that residual, and the reference's own three Heun steps differ from
    it by more than rounding, which the golden shows."""
    say("case: compute_update_exact 300x200 reference build; resident run vs oracle")
    g = load("bench2d_update_exact_300x200")
    p, u, w = bench2d_inputs(salt=int(g["salt"]), nx=300, ny=200)
    with fvwb.Solver2D(fvwb.reference_params(300, 200, 2)) as s:
        d = s.compute_update_exact(u, w)
        c = s.compute_max_speed(u)
    sm = summarize(d)
    assert np.max(np.abs(sm["sample"] - g["sample"])) <= res_tol(300)
    np.testing.assert_allclose(sm["abssum"], g["abssum"], rtol=1e-9)
    assert abs(c - float(g["cmax"][0])) <= 4 * EPS
    ge = load("bench2d_evolve_300x200")
    tend = float(ge["tend"])
    ug, t, n = fvwb.evolve(u, w, fvwb.reference_params(300, 200, 2), t_end=tend)
    uo, to, no, _ = po.fv2d_advance(p, u, w, t_end=tend, max_steps=1000)
    # the device takes the wave speed for dt from the closed-form centre
    # equilibrium plus the deviation, the oracle from the FP32-rounded w_eq of
    # the golden inputs: dt agrees to ~1e-8 relative, the state to ~1e-11
    assert n == no == 3 and abs(t - to) <= 1e-7 * to
    assert np.max(np.abs(ug - uo)) <= 1e-10
    se = summarize(ug - po.conservative(p.gamma, w))
    assert np.max(np.abs(se["sample"] - ge["sample"])) > 1e-9


def test_evolve_matches_oracle_closed_form_equilibrium():
    say("case: evolve (Heun) 300x300, three steps, closed-form w_eq")
    p = po.make_params(300, 300)
    u, w = po.perturbed_state(p, amp_bump=1e-3, amp_noise=1e-4, salt=1)
    c = po.ref2d_max_speed(p, u)
    tend = 2.5 * (0.5 * (1.0 / 300) / c * 0.5)
    ug, t, n = fvwb.evolve(u, w, fvwb.reference_params(300, 300, 2), t_end=tend)
    uo, to, no = po.ref2d_evolve(p, 2, u, w, tend)
    assert n == no == 3 and abs(t - to) <= 1e-14 * to
    assert np.max(np.abs(ug - uo)) <= 1e-12


# ---- every scheme: residual and stepping against the oracle -------------------------
@pytest.mark.parametrize("scheme", SCHEMES)
def test_residual_every_scheme_matches_oracle(scheme):
    r, f, b = scheme
    say(f"case: residual recon={r} flux={f} bc={b} 130x45")
    gp = gpu_params(130, 45, r, f, b, rho0=1.0, g=1.3, p0=1.1)
    op = orc_params(gp)
    u, w = po.perturbed_state(op, amp_bump=1e-2, amp_noise=1e-3, x0=0.6, y0=0.4, salt=3)
    with fvwb.Solver2D(gp) as s:
        d = s.compute_residual(u, w)
    np.testing.assert_allclose(d, po.fv2d_residual(op, u, w), rtol=0, atol=res_tol(130))


@pytest.mark.parametrize("integ", (fvwb.RK2_HEUN, fvwb.SSPRK3))
@pytest.mark.parametrize("scheme", [(1, 1, 1), (0, 0, 0), (1, 0, 1), (0, 1, 0)])
def test_fp64_stepping_matches_oracle(scheme, integ):
    r, f, b = scheme
    say(f"case: fp64 advance recon={r} flux={f} bc={b} integ={integ} 96x80 x 20 steps")
    gp = gpu_params(96, 80, r, f, b, integ, fvwb.FP64, rho0=1.0)
    op = orc_params(gp)
    u, w = po.perturbed_state(op, amp_bump=1e-2, amp_noise=1e-4, salt=4)
    with fvwb.Solver2D(gp) as s:
        s.upload(u, w)
        t, n = s.advance(20)
        ug = s.download()
        _, dt, steps = s.time()
    uo, to, no, dto = po.fv2d_advance(op, u, w, max_steps=20)
    assert n == 20 and steps == 20
    assert abs(t - to) <= 1e-14 * to and abs(dt - dto) <= 1e-14 * dto
    assert np.max(np.abs(ug - uo)) <= 1e-11


@pytest.mark.parametrize("integ", (fvwb.RK2_HEUN, fvwb.SSPRK3))
@pytest.mark.parametrize("scheme", [(1, 1, 1), (0, 0, 0), (1, 0, 0)])
def test_fp32_bulk_stepping_accuracy(scheme, integ):
    r, f, b = scheme
    say(f"case: fp32 advance recon={r} flux={f} bc={b} integ={integ} 128x96 x 30 steps")
    gp = gpu_params(128, 96, r, f, b, integ, fvwb.FP32, rho0=1.0)
    op = orc_params(gp)
    u, w = po.perturbed_state(op, amp_bump=1e-3, amp_noise=0.0, salt=4)
    with fvwb.Solver2D(gp) as s:
        s.upload(u, w)
        t, n = s.advance(30)
        ug = s.download()
    uo, to, no, _ = po.fv2d_advance(op, u, w, max_steps=30)
    err = np.abs(ug[0] - uo[0])
    l1 = err.sum() * (1.0 / 128) * (1.0 / 96)
    say(f"  L1 density error {l1:.3e}, max {err.max():.3e}, t rel diff {abs(t - to) / to:.1e}")
    assert l1 <= 1e-6 and err.max() <= 1e-6
    assert abs(t - to) <= 1e-6 * to


# ---- exact hydrostatic balance -----------------------------------------------------
@pytest.mark.parametrize("prec", (fvwb.FP64, fvwb.FP32))
@pytest.mark.parametrize("scheme", SCHEMES)
def test_resting_atmosphere_is_preserved_bitwise(scheme, prec):
    r, f, b = scheme
    say(f"case: equilibrium 100 steps recon={r} flux={f} bc={b} prec={prec}")
    for atmos in ((fvwb.RHO0_REF, 1.0, 1.0), (1.0, 1.0, 2.0)):
        gp = gpu_params(70, 66, r, f, b, fvwb.SSPRK3, prec, rho0=atmos[0], p0=atmos[1], g=atmos[2])
        w = po.equilibrium_primitive(orc_params(gp))
        u = po.conservative(gp.gamma, w)
        with fvwb.Solver2D(gp) as s:
            s.upload(u, w)
            t, n = s.advance(100)
            dev = s.download_deviation()
            back = s.download()
        assert n == 100 and t > 0
        assert np.count_nonzero(dev) == 0
        assert np.array_equal(back, u)


# ---- shapes ------------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(4, 4), (4, 130), (130, 4), (61, 7), (121, 33), (60, 60), (62, 5)])
@pytest.mark.parametrize("recon", (0, 1))
def test_edge_shapes_residual(shape, recon):
    nx, ny = shape
    say(f"case: residual shape {nx}x{ny} recon={recon}")
    gp = gpu_params(nx, ny, recon, 1, 1, rho0=1.0)
    op = orc_params(gp)
    u, w = po.perturbed_state(op, amp_bump=1e-2, amp_noise=1e-3, salt=6)
    with fvwb.Solver2D(gp) as s:
        d = s.compute_residual(u, w)
    np.testing.assert_allclose(d, po.fv2d_residual(op, u, w), rtol=0, atol=res_tol(max(nx, ny)))


def test_device_initial_condition_matches_host_upload():
    say("case: init_gaussian vs upload 150x150")
    gp = gpu_params(150, 150, 1, 1, 1, fvwb.SSPRK3, fvwb.FP64, rho0=1.0)
    op = orc_params(gp)
    u, w = po.perturbed_state(op, amp_bump=1e-3, amp_noise=0.0)
    with fvwb.Solver2D(gp) as a, fvwb.Solver2D(gp) as b:
        a.upload(u, w)
        b.init_gaussian(1e-3, 0.3, 0.3, 100.0)
        np.testing.assert_allclose(b.download(), u, rtol=0, atol=1e-15)
        a.advance(5)
        b.advance(5)
        np.testing.assert_allclose(b.download(), a.download(), rtol=0, atol=1e-13)


def test_time_bounded_run_stops_at_t_end():
    say("case: advance while t < t_end 64x64")
    gp = gpu_params(64, 64, 1, 1, 1, fvwb.SSPRK3, fvwb.FP64, rho0=1.0)
    op = orc_params(gp)
    u, w = po.perturbed_state(op, amp_bump=1e-3, amp_noise=0.0)
    with fvwb.Solver2D(gp) as s:
        s.upload(u, w)
        t, n = s.advance(100000, t_end=0.05)
    uo, to, no, _ = po.fv2d_advance(op, u, w, t_end=0.05, max_steps=100000)
    assert n == no and abs(t - to) <= 1e-13 and t >= 0.05


def test_nonphysical_state_is_reported():
    say("case: negative density -> FVWB_ERR_NONPHYSICAL")
    gp = gpu_params(64, 64, 1, 1, 1, fvwb.SSPRK3, fvwb.FP64, rho0=1.0)
    op = orc_params(gp)
    u, w = po.perturbed_state(op, amp_bump=1e-3, amp_noise=0.0)
    u[0, 10, 10] = -0.5
    with fvwb.Solver2D(gp) as s:
        s.upload(u, w)
        with pytest.raises(fvwb.FvwbError) as e:
            s.advance(2)
        assert e.value.code == fvwb.ERR_NONPHYSICAL


def test_advance_before_upload_is_a_state_error():
    with fvwb.Solver2D(gpu_params(16, 16)) as s:
        with pytest.raises(fvwb.FvwbError) as e:
            s.advance(1)
        assert e.value.code == fvwb.ERR_STATE


# ---- full benchmark sizes: size-independent properties -------------------------------
def test_2048_fp32_tracks_fp64_and_stays_balanced():
    say("case: 2048x2048 fp32 vs fp64, 20 SSP-RK3 steps; equilibrium 20 steps")
    base = dict(recon=1, flux=1, bc=1, integ=fvwb.SSPRK3, rho0=1.0)
    with fvwb.Solver2D(gpu_params(2048, 2048, prec=fvwb.FP64, **base)) as a, \
            fvwb.Solver2D(gpu_params(2048, 2048, prec=fvwb.FP32, **base)) as b:
        for s in (a, b):
            s.init_gaussian(1e-3, 0.3, 0.3, 100.0)
            s.advance(20)
        da, db = a.download_deviation(), b.download_deviation()
        ta, tb = a.time()[0], b.time()[0]
        l1 = np.abs(da[0] - db[0]).sum() / 2048 ** 2
        say(f"  L1 density fp32 vs fp64 {l1:.3e}, max {np.abs(da[0] - db[0]).max():.3e}")
        assert l1 <= 1e-6 and abs(ta - tb) <= 1e-6 * ta
        # x<->y symmetry of the symmetric bump problem on the full grid
        assert np.max(np.abs(da[0] - da[0].T)) <= 1e-13
        for s in (a, b):
            s.init_gaussian(0.0, 0.3, 0.3, 100.0)
            s.advance(20)
            assert np.count_nonzero(s.download_deviation()) == 0


@pytest.mark.parametrize("precision", [fvwb.FP32, fvwb.FP64])
def test_advance_past_t_end_then_further(precision):
    """Steps past t_end are no-ops that keep the state, the clock and the
    wave-speed maximum, so advancing to t1 and then to t2 ends exactly where
    one advance to t2 ends (same steps, same time, same state)."""
    print(f"case: advance to t1 then t2, precision={precision}", flush=True)
    p = fvwb.make_params(48, 40, bc=fvwb.BC_REFLECT, recon=fvwb.RECON_MUSCL_MINMOD, flux=fvwb.FLUX_HLL,
                         integrator=fvwb.SSPRK3, precision=precision, rho0=1.0)
    with fvwb.Solver2D(p) as a, fvwb.Solver2D(p) as b:
        for s in (a, b):
            s.init_gaussian(1e-2, 0.5, 0.5, 100.0)
        a.advance(1000, 0.01)
        a.advance(1000, 0.025)
        b.advance(2000, 0.025)
        assert a.time() == b.time()
        assert np.array_equal(a.download_deviation(), b.download_deviation())
