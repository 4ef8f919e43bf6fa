"""Batched instances (one persistent launch per stage over every strip of every
instance): each instance must reproduce, bit for bit, what the single-instance
solver computes for it alone (same per-cell arithmetic), with its own dt and
clock; graph replay, launch shape and autotuning must not change results; the
resting atmosphere stays exactly at rest; FP64 instances match the oracle."""
import math
import os
import sys
import tempfile

import numpy as np
import pytest

import fvwb
import pyoracle as po

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


SIZES = [(64, 64), (100, 70), (130, 200), (37, 45), (250, 31), (8, 300)]
GS = [1.0, 0.5, 2.0, 1.3, 0.7, 1.8]


def inst_params(prec, recon=1, flux=1, bc=1, integ=fvwb.SSPRK3, sizes=SIZES, gs=GS):
    return [fvwb.make_params(nx, ny, recon=recon, flux=flux, bc=bc, integrator=integ, precision=prec,
                             gamma=1.4, rho0=1.0, p0=1.0, g=g) for (nx, ny), g in zip(sizes, gs)]


def single_run(p, amp, nsteps, rows=8, cpl=2):
    """The same problem alone, on the batch's strip decomposition (rows, cpl)."""
    with fvwb.Solver2D(p) as s:
        s.set_launch(rows, cpl)
        s.init_gaussian(amp, 0.3, 0.3, 100.0)
        s.advance(nsteps)
        return s.download_deviation(), s.time()


@pytest.mark.parametrize("prec", (fvwb.FP64, fvwb.FP32))
@pytest.mark.parametrize("scheme", [(1, 1, 1), (0, 0, 0)])
def test_batch_equals_single_instances_bitwise(prec, scheme):
    r, f, b = scheme
    say(f"case: batch of {len(SIZES)} vs single instances prec={prec} scheme={scheme}")
    ps = inst_params(prec, r, f, b)
    amps = np.array([1e-3, 1e-2, 5e-3, 1e-4, 2e-3, 7e-3])
    with fvwb.Batch2D(ps) as bt:
        bt.init_gaussian(amps)
        bt.advance(12)
        for k, p in enumerate(ps):
            dev_s, (ts, dts, ns) = single_run(p, amps[k], 12)
            t, dt, n = bt.time(k)
            assert n == ns == 12 and t == ts and dt == dts
            assert np.array_equal(bt.download_deviation(k), dev_s), k


def test_batch_fp64_matches_oracle():
    say("case: batch FP64 instances vs oracle")
    ps = inst_params(fvwb.FP64, sizes=SIZES[:4], gs=GS[:4])
    with fvwb.Batch2D(ps) as bt:
        states = []
        for k, p in enumerate(ps):
            op = po.make_params(p.nx, p.ny, recon=1, flux=1, bc=1, integrator=po.SSPRK3, gamma=1.4, rho0=1.0,
                                p0=1.0, g=p.g)
            u, w = po.perturbed_state(op, amp_bump=1e-3 * (k + 1), amp_noise=1e-4, salt=k + 2)
            bt.upload(k, u, w)
            states.append((op, u, w))
        bt.reset_time()
        bt.advance(10)
        for k, (op, u, w) in enumerate(states):
            uo, to, no, _ = po.fv2d_advance(op, u, w, max_steps=10)
            t, dt, n = bt.time(k)
            assert n == no == 10 and abs(t - to) <= 1e-14 * to
            assert np.max(np.abs(bt.download(k) - uo)) <= 1e-11


def test_graph_launch_shape_and_autotune_do_not_change_results():
    say("case: batch graph / shape / autotune invariance")
    ps = inst_params(fvwb.FP32)
    amps = np.linspace(1e-4, 1e-2, len(ps))
    with fvwb.Batch2D(ps) as a, fvwb.Batch2D(ps) as b:
        a.use_graph(False)
        a.init_gaussian(amps)
        a.advance(7)
        b.set_launch(16, 4, 3)
        b.init_gaussian(amps)
        with tempfile.TemporaryDirectory() as td:
            cache = os.path.join(td, "tune.txt")
            before = [b.download_deviation(k) for k in range(len(ps))]
            rows, cpl = b.autotune(2, cache)
            assert rows % 4 == 0 and cpl in (2, 4)
            for k in range(len(ps)):     # the tuner restores the state
                assert np.array_equal(b.download_deviation(k), before[k])
            assert os.path.exists(cache) and "fv2d-batch" in open(cache).read()
            assert b.autotune(2, cache) == (rows, cpl)   # cached
        b.advance(7)
        same_shape = b.get_launch()[:2] == a.get_launch()[:2]
        for k in range(len(ps)):
            da, db = a.download_deviation(k), b.download_deviation(k)
            if same_shape:
                assert np.array_equal(da, db), k
                assert a.time(k) == b.time(k)
            else:
                # another strip decomposition may round a strip-edge flux differently
                np.testing.assert_allclose(db, da, rtol=0, atol=1e-9)
                assert abs(a.time(k)[0] - b.time(k)[0]) <= 1e-6 * a.time(k)[0]


@pytest.mark.parametrize("prec", (fvwb.FP64, fvwb.FP32))
def test_batch_resting_atmospheres_stay_exactly_at_rest(prec):
    say(f"case: batch equilibrium prec={prec}")
    ps = inst_params(prec)
    with fvwb.Batch2D(ps) as bt:
        bt.init_gaussian(np.zeros(len(ps)))
        bt.advance(40)
        for k in range(len(ps)):
            assert np.count_nonzero(bt.download_deviation(k)) == 0
            assert bt.time(k)[2] == 40


def test_batch_rejects_mixed_schemes_and_bad_indices():
    ps = inst_params(fvwb.FP32, sizes=SIZES[:2], gs=GS[:2])
    ps[1].flux = fvwb.FLUX_RUSANOV
    with pytest.raises(fvwb.FvwbError) as e:
        fvwb.Batch2D(ps)
    assert e.value.code == fvwb.ERR_ARG
    with fvwb.Batch2D(inst_params(fvwb.FP32, sizes=SIZES[:2], gs=GS[:2])) as bt:
        with pytest.raises(fvwb.FvwbError):
            bt.advance(1)           # before init
        bt.init_gaussian(np.array([1e-3, 1e-3]))
        with pytest.raises(fvwb.FvwbError):
            bt.download(5)


def test_config5_like_sample_runs_with_independent_clocks():
    """A 32-instance draw like the 512-instance workload (log-uniform sides
    64..256 here, g ~ U[0.5,2], amplitude ~ U[1e-4,1e-2]): every instance
    advances exactly 20 steps with its own dt = 0.5*h/cmax*cfl."""
    say("case: 32-instance log-uniform batch")
    rng = np.random.default_rng(20261202)
    sides = np.rint(np.exp(rng.uniform(np.log(64), np.log(256), 32))).astype(int)
    gs = rng.uniform(0.5, 2.0, 32)
    amps = rng.uniform(1e-4, 1e-2, 32)
    ps = [fvwb.make_params(int(n), int(n), recon=1, flux=1, bc=1, integrator=fvwb.SSPRK3, precision=fvwb.FP32,
                           gamma=1.4, rho0=1.0, p0=1.0, g=float(g)) for n, g in zip(sides, gs)]
    with fvwb.Batch2D(ps) as bt:
        bt.init_gaussian(amps)
        bt.advance(20)
        dts = []
        for k in (0, 7, 31):
            t, dt, n = bt.time(k)
            assert n == 20 and t > 0
            dts.append(dt)
            dev_s, (ts, dts_s, ns) = single_run(ps[k], amps[k], 20)
            assert np.array_equal(bt.download_deviation(k), dev_s) and t == ts
        assert len(set(dts)) == 3
