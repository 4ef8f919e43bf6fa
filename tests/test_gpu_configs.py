"""The two FP32-bulk workloads at their full sizes.

- 8192 x 8192 gravity problem (MUSCL-minmod, HLL, reflecting walls, SSP-RK3,
  rho = p = exp(-(x+y)) plus a pressure bump 1e-3 exp(-100 r^2) at (0.3, 0.3)):
  FP32-bulk deviation form against the plain FP64 kernel, the x<->y symmetry of
  the problem, and a pure-equilibrium run that must stay exactly at rest.
- 512 independent instances, sides log-uniform in [64, 1024] (seed 20261202),
  g ~ U[0.5, 2], bump amplitude ~ U[1e-4, 1e-2]: the FP32 batch against the
  FP64 batch on every instance, and four of the smallest instances against the
  FP64 oracle.

Accuracy: L1 = sum |rho - rho_ref| dx dy over the unit square (the density
error the precision target is stated in, bound 1e-6); L_inf is reported.
FP64 against the oracle: |du| <= 1e-11 after 5 steps (FMA contraction and the
separable equilibrium tables, a few ulps per stage)."""
import sys

import numpy as np
import pytest

import fvwb
import pyoracle as po

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def scheme(nx, ny, precision, g=1.0):
    return fvwb.make_params(nx, ny, recon=fvwb.RECON_MUSCL_MINMOD, flux=fvwb.FLUX_HLL, bc=fvwb.BC_REFLECT,
                            integrator=fvwb.SSPRK3, precision=precision, gamma=1.4, rho0=1.0, p0=1.0, g=g)


def batch_workload(n=512, seed=20261202):
    """The 512-instance draw, in the order the bench draws it."""
    rng = np.random.default_rng(seed)
    sides = np.rint(np.exp(rng.uniform(np.log(64), np.log(1024), n))).astype(int)
    gs = rng.uniform(0.5, 2.0, n)
    amps = rng.uniform(1e-4, 1e-2, n)
    return sides, gs, amps


def test_8192_fp32_bulk_against_fp64():
    n, steps = 8192, 200
    say(f"case: {n}x{n} fp32-bulk vs fp64, {steps} SSP-RK3 steps")
    with fvwb.Solver2D(scheme(n, n, fvwb.FP64)) as a:
        a.init_gaussian(1e-3, 0.3, 0.3, 100.0)
        a.advance(steps)
        ta = a.time()[0]
        r64 = a.download_deviation()[0]
    with fvwb.Solver2D(scheme(n, n, fvwb.FP32)) as b:
        b.init_gaussian(1e-3, 0.3, 0.3, 100.0)
        b.advance(steps)
        tb = b.time()[0]
        r32 = b.download_deviation()[0]
    diff = np.abs(r32 - r64)
    l1, linf = float(diff.sum()) / n ** 2, float(diff.max())
    say(f"  density fp32 vs fp64: L1 {l1:.3e}, Linf {linf:.3e}; max |drho| {np.abs(r64).max():.3e}; "
        f"t {ta:.6e} vs {tb:.6e}")
    assert l1 <= 1e-6 and linf <= 1e-6
    assert abs(ta - tb) <= 1e-6 * ta
    assert np.abs(r64).max() > 1e-8          # the density has responded: the comparison is not of zeros
    # the problem is symmetric under x <-> y (phi = x + y, bump on the diagonal)
    assert np.max(np.abs(r64 - r64.T)) <= 1e-13
    assert np.max(np.abs(r32 - r32.T)) <= 1e-9


@pytest.mark.parametrize("precision,steps", [(fvwb.FP32, 2000), (fvwb.FP64, 200)])
def test_8192_resting_atmosphere_stays_exactly_at_rest(precision, steps):
    n = 8192
    say(f"case: {n}x{n} pure equilibrium, precision={precision}, {steps} steps")
    with fvwb.Solver2D(scheme(n, n, precision)) as s:
        s.init_gaussian(0.0, 0.3, 0.3, 100.0)
        s.advance(steps)
        t, dt, k = s.time()
        dev = s.download_deviation()
    assert k == steps and t > 0 and dt > 0
    nz = int(np.count_nonzero(dev))
    say(f"  drift: max |drho| {np.abs(dev[0]).max():.1e}, max |dE| {np.abs(dev[3]).max():.1e}, nonzero {nz}")
    assert nz == 0


def test_512_instance_batch_fp32_against_fp64_and_oracle():
    sides, gs, amps = batch_workload()
    cells = int(np.sum(sides.astype(np.int64) ** 2))
    steps = 5
    say(f"case: 512-instance batch ({cells} cells, sides {sides.min()}..{sides.max()}), {steps} steps, "
        f"fp32 vs fp64 on every instance, 4 instances vs the oracle")
    small = [int(k) for k in np.argsort(sides, kind="stable")[:4]]
    with fvwb.Batch2D([scheme(int(s), int(s), fvwb.FP64, float(g)) for s, g in zip(sides, gs)]) as b64:
        b64.init_gaussian(amps)
        u0 = {k: b64.download(k) for k in small}
        b64.advance(steps)
        r64 = [b64.download_deviation(k)[0] for k in range(len(sides))]
        u64 = {k: b64.download(k) for k in small}
        clocks64 = [b64.time(k) for k in range(len(sides))]
    with fvwb.Batch2D([scheme(int(s), int(s), fvwb.FP32, float(g)) for s, g in zip(sides, gs)]) as b32:
        b32.init_gaussian(amps)
        b32.advance(steps)
        u32 = {k: b32.download(k) for k in small}
        l1s, linfs = [], []
        for k, s in enumerate(sides):
            d = np.abs(b32.download_deviation(k)[0] - r64[k])
            l1s.append(float(d.sum()) / (int(s) ** 2))
            linfs.append(float(d.max()))
            t32, _, n32 = b32.time(k)
            assert n32 == clocks64[k][2] == steps
            assert abs(t32 - clocks64[k][0]) <= 1e-6 * clocks64[k][0]
    k1 = int(np.argmax(l1s))
    say(f"  fp32 vs fp64 density: max L1 {max(l1s):.3e} (instance {k1}, side {sides[k1]}), "
        f"max Linf {max(linfs):.3e}, mean L1 {np.mean(l1s):.3e}")
    assert max(l1s) <= 1e-6 and max(linfs) <= 1e-6
    for k in small:
        s = int(sides[k])
        op = po.make_params(s, s, recon=po.RECON_MUSCL_MINMOD, flux=po.FLUX_HLL, bc=po.BC_REFLECT,
                            integrator=po.SSPRK3, gamma=1.4, rho0=1.0, p0=1.0, g=float(gs[k]))
        w = po.equilibrium_primitive(op)
        uo, to, no, _ = po.fv2d_advance(op, u0[k], w, max_steps=steps)
        err = float(np.max(np.abs(u64[k] - uo)))
        say(f"  instance {k} (side {s}, g {gs[k]:.3f}, amp {amps[k]:.2e}): fp64 batch vs oracle max {err:.2e}, "
            f"t {clocks64[k][0]:.6e} vs {to:.6e}")
        l1_32 = float(np.abs(u32[k][0] - uo[0]).sum()) / s ** 2
        say(f"    fp32 batch vs oracle: density L1 {l1_32:.3e}, Linf {np.abs(u32[k][0] - uo[0]).max():.3e}")
        assert no == steps and abs(clocks64[k][0] - to) <= 1e-13 * to
        assert err <= 1e-11 and l1_32 <= 1e-6
