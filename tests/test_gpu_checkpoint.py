"""Checkpoint / resume of every device-resident time loop: a handle that loads
a checkpoint continues bit for bit as the handle that saved it, and a load
rejects a foreign, mismatched or damaged file."""
import os
import sys

import numpy as np
import pytest

import fvwb
import pyoracle as po

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def resume_equals_run(make, start, run, read, path, n0, n1):
    """start a handle, run n0 steps, save, run n1 more; a second handle loads
    the file and runs n1 steps; both read the same, bit for bit."""
    with make() as a:
        start(a)
        run(a, n0)
        a.checkpoint_save(path)
        run(a, n1)
        ra = read(a)
    with make() as b:
        b.checkpoint_load(path)
        run(b, n1)
        rb = read(b)
    for x, y in zip(ra, rb):
        assert np.array_equal(np.asarray(x), np.asarray(y))
    return ra


@pytest.mark.parametrize("precision", [fvwb.FP32, fvwb.FP64])
def test_fv2d_resume(tmp_path, precision):
    say(f"case: 2D FV checkpoint precision={precision}")
    p = fvwb.make_params(64, 48, bc=fvwb.BC_REFLECT, recon=fvwb.RECON_MUSCL_MINMOD, flux=fvwb.FLUX_HLL,
                         integrator=fvwb.SSPRK3, precision=precision, rho0=1.0)
    resume_equals_run(lambda: fvwb.Solver2D(p), lambda s: s.init_gaussian(1e-2, 0.5, 0.5, 100.0),
                      lambda s, n: s.advance(n), lambda s: (s.download_deviation(), s.time()),
                      str(tmp_path / "fv2d.ckpt"), 17, 23)


def test_fv2d_reference_setup_resume(tmp_path):
    say("case: 2D FV checkpoint, the reference set-up (frozen ring, Heun)")
    p = fvwb.reference_params(40, 40, 2)
    u, w = po.perturbed_state(po.make_params(40, 40), amp_bump=1e-3, amp_noise=1e-4, salt=3)
    resume_equals_run(lambda: fvwb.Solver2D(p), lambda s: s.upload(u, w), lambda s, n: s.advance(n),
                      lambda s: (s.download(), s.time()), str(tmp_path / "ref.ckpt"), 5, 9)


def test_fv1d_resume(tmp_path):
    say("case: 1D FV checkpoint, three uneven instances")
    ps = [fvwb.make_params_1d(n, recon=fvwb.RECON_MUSCL_MINMOD) for n in (64, 100, 37)]
    def read(s):
        return [s.download(b) for b in range(3)] + [s.time(b) for b in range(3)]
    resume_equals_run(lambda: fvwb.Solver1D(ps), lambda s: s.init_gaussian([1e-3, 2e-3, 5e-4]),
                      lambda s, n: s.advance(n), read, str(tmp_path / "fv1d.ckpt"), 30, 41)


@pytest.mark.parametrize("limiter", [po.DG_LIMITER_TVB, po.DG_LIMITER_TVB_POSITIVITY])
def test_dg2d_resume(tmp_path, limiter):
    say(f"case: 2D DG checkpoint limiter={limiter}")
    op = po.make_dg_params(40, 36, limiter=limiter, integrator=po.DG_SSPRK54)
    gp = fvwb.DGParams(op.nvar, op.nx, op.ny, op.degree, op.flux, op.limiter, op.bc, op.integrator, op.gamma,
                       op.boxlen_x, op.boxlen_y, op.cfl, op.rho0, op.p0, op.g, op.tvb_m)
    d = po.dg2d_fourier_state(op, amp=1e-3)
    resume_equals_run(lambda: fvwb.SolverDG(gp), lambda s: s.upload(d), lambda s, n: s.advance(n),
                      lambda s: (s.download(), s.time()), str(tmp_path / "dg.ckpt"), 7, 11)


def test_dg1d_resume(tmp_path):
    say("case: 1D DG checkpoint, the reference set-up")
    op = po.make_dg1d_params(128)
    d0, e, u0 = po.dg1d_initial(op, ninit=8, pert=1e-3)
    resume_equals_run(lambda: fvwb.SolverDG1D(fvwb.reference_params_dg1d(128)), lambda s: s.upload(d0, e, u0),
                      lambda s, n: s.advance(n), lambda s: s.download() + (s.time(),), str(tmp_path / "dg1d.ckpt"),
                      50, 60)


def test_rd_resume(tmp_path):
    say("case: residual-distribution checkpoint, 65x65 nodes, ninit 8")
    op = po.make_rd_params(65, 65, bc=po.RD_BC_TRANSMISSIVE)
    u0 = po.rd_initial(op, 8, 1e-2)
    resume_equals_run(lambda: fvwb.SolverRD(fvwb.make_rd_params(65, 65)), lambda s: s.upload(u0),
                      lambda s, n: s.advance(n), lambda s: (s.download(), s.time()), str(tmp_path / "rd.ckpt"),
                      40, 30)


def test_load_rejects_foreign_and_damaged_files(tmp_path):
    say("case: checkpoint load errors")
    p = fvwb.make_params(32, 32, bc=fvwb.BC_REFLECT, precision=fvwb.FP32, rho0=1.0)
    path = str(tmp_path / "a.ckpt")
    with fvwb.Solver2D(p) as s:
        with pytest.raises(fvwb.FvwbError, match="upload"):
            s.checkpoint_save(path)
        s.init_gaussian(1e-3, 0.5, 0.5, 100.0)
        s.advance(3)
        s.checkpoint_save(path)
    q = fvwb.make_params(32, 32, bc=fvwb.BC_REFLECT, precision=fvwb.FP32, rho0=1.0, g=2.0)
    with fvwb.Solver2D(q) as s:
        with pytest.raises(fvwb.FvwbError, match="other parameters"):
            s.checkpoint_load(path)
        with pytest.raises(fvwb.FvwbError):
            s.advance(1)                                   # no state after a failed load
    with fvwb.SolverDG1D(fvwb.reference_params_dg1d(32)) as s:
        with pytest.raises(fvwb.FvwbError, match="another handle type"):
            s.checkpoint_load(path)
    data = open(path, "rb").read()
    assert data[:8] == b"FVWBCKPT" and int.from_bytes(data[8:12], "little") == 2
    cut, v1 = str(tmp_path / "cut.ckpt"), str(tmp_path / "v1.ckpt")
    with open(cut, "wb") as f:
        f.write(data[: len(data) // 2])
    # version 1 held the 2D state as separate planes: same sizes, other layout
    with open(v1, "wb") as f:
        f.write(data[:8] + (1).to_bytes(4, "little") + data[12:])
    with fvwb.Solver2D(p) as s:
        with pytest.raises(fvwb.FvwbError, match="version"):
            s.checkpoint_load(v1)
        with pytest.raises(fvwb.FvwbError, match="truncated"):
            s.checkpoint_load(cut)
        with pytest.raises(fvwb.FvwbError, match="cannot open"):
            s.checkpoint_load(str(tmp_path / "missing.ckpt"))


def test_failed_load_into_a_live_handle(tmp_path):
    """A handle that holds state and then fails to load a truncated or padded
    file is left without state (its next advance is refused), and the file is
    rejected before any of it reaches the device: a good load afterwards
    resumes bit for bit."""
    say("case: failed checkpoint load into a handle with state")
    p = fvwb.make_params(48, 40, bc=fvwb.BC_REFLECT, recon=fvwb.RECON_MUSCL_MINMOD, flux=fvwb.FLUX_HLL,
                         integrator=fvwb.SSPRK3, precision=fvwb.FP32, rho0=1.0)
    path = str(tmp_path / "good.ckpt")
    with fvwb.Solver2D(p) as a:
        a.init_gaussian(1e-2, 0.4, 0.6, 80.0)
        a.advance(6)
        a.checkpoint_save(path)
        a.advance(5)
        want = a.download_deviation()
    data = open(path, "rb").read()
    cut, padded = str(tmp_path / "cut.ckpt"), str(tmp_path / "padded.ckpt")
    with open(cut, "wb") as f:
        f.write(data[:-100])                       # inside the last segment
    with open(padded, "wb") as f:
        f.write(data + b"\0" * 8)
    for bad, why in ((cut, "truncated"), (padded, "trailing bytes")):
        with fvwb.Solver2D(p) as b:
            b.init_gaussian(1e-3, 0.5, 0.5, 100.0)
            b.advance(2)
            with pytest.raises(fvwb.FvwbError, match=why):
                b.checkpoint_load(bad)
            with pytest.raises(fvwb.FvwbError, match="upload"):
                b.advance(1)
            b.checkpoint_load(path)
            b.advance(5)
            assert np.array_equal(b.download_deviation(), want)
    for make, bad in ((lambda: fvwb.SolverRD(fvwb.make_rd_params(33, 33)), cut),):
        op = po.make_rd_params(33, 33, bc=po.RD_BC_TRANSMISSIVE)
        rpath = str(tmp_path / "rd.ckpt")
        with make() as r:
            r.upload(po.rd_initial(op, 8, 1e-2))
            r.advance(3)
            r.checkpoint_save(rpath)
        rdata = open(rpath, "rb").read()
        with open(bad, "wb") as f:
            f.write(rdata[:-8])
        with make() as r:
            r.upload(po.rd_initial(op, 8, 1e-2))
            with pytest.raises(fvwb.FvwbError, match="truncated"):
                r.checkpoint_load(bad)
            with pytest.raises(fvwb.FvwbError):
                r.advance(1)
