"""The Fortran side of the boundary: the iso_c_binding interface module, the
drop-in compute_update_exact / compute_max_speed (benchmark_2d.f90:465-618,
:264-279) and the 2D benchmark driver built on them.

CPU: the Fortran view of fvwb_params (size, field order, the reference set-up
values) equals the C view, the error path returns FVWB_ERR_ARG with a message,
and the driver binds the drop-in symbols to the library.  GPU: the driver's
reference-style Heun loop through the drop-ins reproduces the oracle's
restatement of the reference's evolve on the driver's own initial state; the
device-resident mode agrees with it; the reference set-up (pure equilibrium)
stays bitwise at rest; snapshots are in the reference's ASCII format."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import fvwb
import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "fvm-source-wb_amd", "bin")


def exe(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (run __graft_entry__.build())")
    return path


def run(args, **kw):
    print("run:", " ".join(args), flush=True)
    return subprocess.run(args, capture_output=True, text=True, timeout=300, **kw)


def read_state(path):
    with open(path, "rb") as f:
        nx, ny, nvar, steps = np.frombuffer(f.read(16), dtype=np.int32)
        t = float(np.frombuffer(f.read(8), dtype=np.float64)[0])
        n = int(nvar) * int(nx) * int(ny)
        arrs = np.frombuffer(f.read(3 * n * 8), dtype=np.float64)
    shape = (int(nvar), int(nx), int(ny))
    u0, w_eq, u = (arrs[k * n:(k + 1) * n].reshape(shape, order="F") for k in range(3))
    return int(steps), t, u0, w_eq, u


# ---- CPU ------------------------------------------------------------------------------
def test_fortran_params_struct_matches_c_view():
    r = run([exe("fvwb_abi_check")])
    assert r.returncode == 0, r.stderr
    lines = dict(line.split(" ", 1) for line in r.stdout.strip().splitlines())
    assert int(lines["sizeof"]) == ctypes.sizeof(fvwb.Params) == 88
    assert int(lines["status"]) == fvwb.OK
    c = fvwb.reference_params(300, 200, 2)
    ints = [int(v) for v in lines["ints"].split()]
    assert ints == [c.nvar, c.nx, c.ny, c.bc, c.recon, c.flux, c.integrator, c.precision]
    doubles = [np.frombuffer(bytes.fromhex(h), dtype=">f8")[0] for h in lines["doubles"].split()]
    assert doubles == [c.gamma, c.boxlen_x, c.boxlen_y, c.cfl, c.rho0, c.p0, c.g]
    assert doubles[0] == float(np.float32(1.4)) and doubles[4] == float(np.float32(1.21))
    assert int(lines["create_bad"]) == fvwb.ERR_ARG
    assert "nx" in lines["message"]
    assert int(lines["sizeof1d"]) == ctypes.sizeof(fvwb.Params1D) == 56
    c1 = fvwb.reference_params_1d(128, 2)
    assert [int(v) for v in lines["ints1d"].split()] == [c1.nvar, c1.nx, c1.bc, c1.recon, c1.integrator,
                                                          c1.nequilibrium]
    d1 = [np.frombuffer(bytes.fromhex(h), dtype=">f8")[0] for h in lines["doubles1d"].split()]
    assert d1 == [c1.gamma, c1.boxlen, c1.cfl, c1.g]
    assert int(lines["sizeofdg"]) == ctypes.sizeof(fvwb.DGParams) == 96
    assert int(lines["sizeofdgq"]) == ctypes.sizeof(fvwb.DGQParams) == 80
    q = fvwb.reference_params_dgq()
    assert [int(v) for v in lines["intsdgq"].split()] == [q.nx, q.ny, q.order, q.flux_type, q.source,
                                                          q.grad_phi_case, q.bc, q.ninit, q.solver, q.use_limiter]
    dq = [np.frombuffer(bytes.fromhex(h), dtype=">f8")[0] for h in lines["doublesdgq"].split()]
    assert dq == [q.gamma, q.boxlen_x, q.boxlen_y, q.cfl, q.eps]
    assert int(lines["createdgq_bad"]) == fvwb.ERR_ARG
    g3 = [np.frombuffer(bytes.fromhex(h), dtype=">f8")[0] for h in lines["glq3"].split()]
    x3, w3 = fvwb.dgq_gl_quadrature(3)
    assert g3 == list(x3) + list(w3)


def test_driver_binds_dropins_to_library():
    path = exe("benchmark_2d_fvwb")
    nm = subprocess.run(["nm", path], capture_output=True, text=True, check=True).stdout
    defined = set(re.findall(r"^[0-9a-f]+ T (\S+)$", nm, re.M))
    undefined = set(re.findall(r"^\s+U (\S+)$", nm, re.M))
    # the external procedures a reference-style caller binds
    assert {"compute_update_exact_", "compute_max_speed_"} <= defined
    for sym in ("fvwb_compute_update_exact", "fvwb_compute_max_speed", "fvwb_create", "fvwb_upload",
                "fvwb_advance", "fvwb_download", "fvwb_last_error"):
        assert sym in undefined and sym in fvwb.declared_symbols()


def test_1d_driver_binds_dropins_to_library():
    nm = subprocess.run(["nm", exe("benchmark_1d_fvwb")], capture_output=True, text=True, check=True).stdout
    defined = set(re.findall(r"^[0-9a-f]+ T (\S+)$", nm, re.M))
    undefined = set(re.findall(r"^\s+U (\S+)$", nm, re.M))
    assert {"compute_update_", "compute_update_sr_", "compute_max_speed_"} <= defined
    for sym in ("fvwb1d_compute_update", "fvwb1d_compute_update_sr", "fvwb1d_compute_max_speed",
                "fvwb1d_create", "fvwb1d_advance"):
        assert sym in undefined and sym in fvwb.declared_symbols()


def test_fvm_driver_binds_dropins_to_library():
    nm = subprocess.run(["nm", exe("fvm_fvwb")], capture_output=True, text=True, check=True).stdout
    defined = set(re.findall(r"^[0-9a-f]+ T (\S+)$", nm, re.M))
    undefined = set(re.findall(r"^\s+U (\S+)$", nm, re.M))
    assert {"compute_update_", "compute_max_speed_"} <= defined
    for sym in ("fvwb1d_fvm_compute_update", "fvwb1d_compute_max_speed", "fvwb1d_create"):
        assert sym in undefined and sym in fvwb.declared_symbols()


def test_dg1d_driver_binds_dropin_to_library():
    nm = subprocess.run(["nm", exe("dg1d_fvwb")], capture_output=True, text=True, check=True).stdout
    defined = set(re.findall(r"^[0-9a-f]+ T (\S+)$", nm, re.M))
    undefined = set(re.findall(r"^\s+U (\S+)$", nm, re.M))
    assert "compute_update_exact_delta_" in defined
    for sym in ("fvwb_dg1d_compute_update_exact_delta", "fvwb_dg1d_create", "fvwb_dg1d_upload",
                "fvwb_dg1d_advance"):
        assert sym in undefined and sym in fvwb.declared_symbols()


def test_dg2d_driver_binds_dropins_to_library():
    """The reference's 2D DG program binds its external procedures to the GPU
    drop-ins, which call the fvwb_dgq_* entry points the header declares."""
    nm = subprocess.run(["nm", exe("benchmark_2d_dg_fvwb")], capture_output=True, text=True, check=True).stdout
    defined = set(re.findall(r"^[0-9a-f]+ T (\S+)$", nm, re.M))
    undefined = set(re.findall(r"^\s+U (\S+)$", nm, re.M))
    assert {"compute_update_", "compute_max_speed_", "apply_limiter_", "get_modes_from_nodes_",
            "get_nodes_from_modes_", "gl_quadrature_", "compute_error_", "evolve_"} <= defined
    for sym in ("fvwb_dgq_compute_update", "fvwb_dgq_compute_max_speed", "fvwb_dgq_apply_limiter",
                "fvwb_dgq_modes_from_nodes", "fvwb_dgq_nodes_from_modes", "fvwb_dgq_gl_quadrature",
                "fvwb_dgq_upload", "fvwb_dgq_advance", "fvwb_dgq_compute_error"):
        assert sym in undefined and sym in fvwb.declared_symbols()


def test_rd_driver_binds_dropin_to_library():
    nm = subprocess.run(["nm", exe("rd_fvwb")], capture_output=True, text=True, check=True).stdout
    defined = set(re.findall(r"^[0-9a-f]+ T (\S+)$", nm, re.M))
    undefined = set(re.findall(r"^\s+U (\S+)$", nm, re.M))
    assert "compute_update_rd_" in defined and "compute_max_speed_" in defined
    for sym in ("fvwb_rd_compute_update", "fvwb_rd_compute_max_speed", "fvwb_rd_create", "fvwb_rd_upload",
                "fvwb_rd_advance"):
        assert sym in undefined and sym in fvwb.declared_symbols()


def test_driver_without_device_fails_loudly():
    """Without a usable GPU the driver stops with the library's error instead
    of computing anything on the host."""
    try:
        import torch
        if torch.cuda.device_count() > 0:
            pytest.skip("a GPU is visible")
    except ImportError:
        pass
    r = run([exe("benchmark_2d_fvwb_64x48"), "stage", "0.01"])
    assert r.returncode != 0
    assert "fvwb: fvwb_create failed" in r.stderr


# ---- GPU ------------------------------------------------------------------------------
@pytest.mark.gpu
def test_driver_stage_loop_matches_oracle_evolve(tmp_path):
    """64x48, ninit=3 (equilibrium + pulse of amplitude eta), reference set-up
    otherwise: the host Heun loop over the GPU drop-ins reproduces the oracle's
    restatement of benchmark_2d.f90:221-260 to 1e-12 with the same step count."""
    tend = 0.03
    prefix = str(tmp_path / "s")
    r = run([exe("benchmark_2d_fvwb_64x48"), "stage", repr(tend), prefix])
    assert r.returncode == 0, r.stdout + r.stderr
    steps, t, u0, w_eq, u = read_state(prefix + ".bin")
    p = po.make_params(64, 48)
    uo, to, no = po.ref2d_evolve(p, 2, u0, w_eq, tend)
    print(f"steps {steps} vs {no}, t {t} vs {to}, max|du| {np.max(np.abs(u - uo)):.3e}", flush=True)
    assert steps == no > 5 and abs(t - to) <= 1e-13 * to
    assert np.max(np.abs(u - uo)) <= 1e-12
    assert np.max(np.abs(u - u0)) > 1e-7        # the pulse moved


@pytest.mark.gpu
def test_driver_resident_mode_matches_oracle_advance(tmp_path):
    """Device-resident mode integrates the configured residual (y faces at
    (j-1)*dy, which differs from the reference's (j-1)*dx on this 64x48 grid)
    with the same Heun loop: it matches the oracle's configured advance, and it
    stays within the face-position difference of the stage-by-stage run."""
    tend = 0.03
    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    for mode, prefix in (("stage", a), ("resident", b)):
        r = run([exe("benchmark_2d_fvwb_64x48"), mode, repr(tend), prefix])
        assert r.returncode == 0, r.stdout + r.stderr
    sa, ta, _, _, ua = read_state(a + ".bin")
    sb, tb, u0, w_eq, ub = read_state(b + ".bin")
    uo, to, no, _ = po.fv2d_advance(po.make_params(64, 48), u0, w_eq, t_end=tend, max_steps=1 << 20)
    print(f"resident: steps {sb} vs {no}, t {tb} vs {to}, max|du| {np.max(np.abs(ub - uo)):.3e}; "
          f"vs stage {np.max(np.abs(ub - ua)):.3e}", flush=True)
    assert sb == no == sa and abs(tb - to) <= 1e-13 * to
    assert np.max(np.abs(ub - uo)) <= 1e-12
    assert np.max(np.abs(ub - ua)) <= 1e-8


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ("stage", "resident"))
def test_reference_setup_stays_exactly_at_rest(tmp_path, mode):
    """The reference set-up (300x300, ninit=2: the atmosphere itself) through
    the drop-ins: every step returns the initial state bit for bit, and the
    snapshot of p - p_eq, in the reference's format, is zero to rounding."""
    prefix = str(tmp_path / mode)
    r = run([exe("benchmark_2d_fvwb"), mode, "0.02", prefix])
    assert r.returncode == 0, r.stdout + r.stderr
    steps, t, u0, w_eq, u = read_state(prefix + ".bin")
    assert steps >= 5 and np.array_equal(u, u0)
    lines = open(prefix + "_final.dat").read().splitlines()
    assert len(lines) == 300 * 300
    # '(7(1PE12.5,1X))': fields of 12 characters separated by one blank
    fmt = re.compile(r"(?:[ -]\d\.\d{5}E[+-]\d\d ?){3}")
    assert all(fmt.fullmatch(ln) for ln in lines)
    # p recomputed from the (unchanged) conservative state: a rounding-level
    # round trip (gamma-1)*(p/(gamma-1)), as the reference's own snapshot shows
    vals = np.array([float(ln.split()[2]) for ln in lines])
    assert np.max(np.abs(vals)) <= 4e-16
    x0, y0 = (float(v) for v in lines[1].split()[:2])
    assert abs(x0 - 0.5 / 300) < 1e-8 and abs(y0 - 1.5 / 300) < 1e-8
    # the reader's contract (post_processing2d.py:10-14): C-order reshape to
    # (nx, ny) puts x along the first axis and y along the second
    data = np.loadtxt(prefix + "_final.dat")
    X, Y = data[:, 0].reshape((300, 300), order="C"), data[:, 1].reshape((300, 300), order="C")
    assert np.all(np.diff(X, axis=0) > 0) and np.all(np.diff(X, axis=1) == 0)
    assert np.all(np.diff(Y, axis=1) > 0) and np.all(np.diff(Y, axis=0) == 0)


def read_state_1d(path):
    with open(path, "rb") as f:
        nx, nvar, steps = np.frombuffer(f.read(12), dtype=np.int32)
        t = float(np.frombuffer(f.read(8), dtype=np.float64)[0])
        n = int(nvar) * int(nx)
        arrs = np.frombuffer(f.read(3 * n * 8), dtype=np.float64)
    shape = (int(nvar), int(nx))
    u0, w_eq, u = (arrs[k * n:(k + 1) * n].reshape(shape, order="F") for k in range(3))
    return int(steps), t, u0, w_eq, u


@pytest.mark.gpu
def test_1d_driver_stage_loop_matches_oracle_bitwise(tmp_path):
    """nx=64, ninit=2: the driver's Heun loop over the GPU drop-ins equals the
    same loop over the oracle's restatement of benchmark_1d.f90, bit for bit
    (the drop-ins and the oracle round identically, and the loop arithmetic is
    elementwise without contraction on both sides)."""
    tend = 0.05
    prefix = str(tmp_path / "s1")
    r = run([exe("benchmark_1d_fvwb_64"), "stage", repr(tend), prefix])
    assert r.returncode == 0, r.stdout + r.stderr
    steps, t, u0, w_eq, u = read_state_1d(prefix + ".bin")
    p = po.make_params_1d(64)
    dx = 1.0 / 64
    uo, to, n = np.asfortranarray(u0.copy()), 0.0, 0
    while to < tend:
        cmax = po.ref1d_max_speed(p, uo)
        dt = float(np.float32(0.8)) * dx / cmax / (2.0 * 1.0 + 1.0)
        d = po.ref1d_update(p, uo, w_eq)
        w1 = uo + dt * d
        d = po.ref1d_update(p, w1, w_eq)
        uo = 0.5 * uo + 0.5 * w1 + 0.5 * dt * d
        to, n = to + dt, n + 1
    print(f"steps {steps} vs {n}, t {t} vs {to}, max|du| {np.max(np.abs(u - uo)):.3e}", flush=True)
    assert steps == n and t == to
    assert np.array_equal(u, uo)


@pytest.mark.gpu
def test_1d_driver_wb1_loop_matches_oracle(tmp_path):
    """nx=64, ninit=2, the 'WB1' branch of evolve (benchmark_1d.f90:240-246):
    the Heun loop over the drop-in compute_update_sr follows the same loop over
    the oracle's restatement with the same step count; the device pow rounds
    differently now and then, so the states agree to 1e-12, not bit for bit."""
    tend = 0.05
    prefix = str(tmp_path / "w1")
    r = run([exe("benchmark_1d_fvwb_64"), "wb1", repr(tend), prefix])
    assert r.returncode == 0, r.stdout + r.stderr
    steps, t, u0, w_eq, u = read_state_1d(prefix + ".bin")
    p = po.make_params_1d(64)
    dx = 1.0 / 64
    uo, to, n = np.asfortranarray(u0.copy()), 0.0, 0
    while to < tend:
        cmax = po.ref1d_max_speed(p, uo)
        dt = float(np.float32(0.8)) * dx / cmax / (2.0 * 1.0 + 1.0)
        d = po.ref1d_update_sr(p, uo)
        w1 = uo + dt * d
        d = po.ref1d_update_sr(p, w1)
        uo = 0.5 * uo + 0.5 * w1 + 0.5 * dt * d
        to, n = to + dt, n + 1
    print(f"steps {steps} vs {n}, t {t} vs {to}, max|du| {np.max(np.abs(u - uo)):.3e}", flush=True)
    assert steps == n and abs(t - to) <= 1e-14
    assert np.max(np.abs(u - uo)) <= 1e-12


def oracle_heun_1d(residual, u0, tend, dx, stages_per_dt):
    """The drivers' loop: dt = 0.8*dx/cmax/(2n+1) with the FP32 literal 0.8, Heun."""
    p = po.make_params_1d(u0.shape[1])
    uo, to, n = np.asfortranarray(u0.copy()), 0.0, 0
    while to < tend:
        cmax = po.ref1d_max_speed(p, uo)
        dt = float(np.float32(0.8)) * dx / cmax / stages_per_dt
        d = residual(uo)
        w1 = uo + dt * d
        d = residual(w1)
        uo = 0.5 * uo + 0.5 * w1 + 0.5 * dt * d
        to, n = to + dt, n + 1
    return uo, to, n


@pytest.mark.gpu
def test_1d_driver_fvm_loop_matches_oracle_bitwise(tmp_path):
    """nx=64, the plain 'FVM' residual of benchmark_1d.f90 through the drop-in
    compute_update_fvm: the same loop over the oracle, bit for bit."""
    prefix = str(tmp_path / "f1")
    r = run([exe("benchmark_1d_fvwb_64"), "fvm", "0.05", prefix])
    assert r.returncode == 0, r.stdout + r.stderr
    steps, t, u0, w_eq, u = read_state_1d(prefix + ".bin")
    p = po.make_params_1d(64)
    uo, to, n = oracle_heun_1d(lambda v: po.ref1d_update_fvm(p, v), u0, 0.05, 1.0 / 64, 3.0)
    print(f"steps {steps} vs {n}, t {t} vs {to}, max|du| {np.max(np.abs(u - uo)):.3e}", flush=True)
    assert steps == n and t == to
    assert np.array_equal(u, uo)


@pytest.mark.gpu
def test_fvm_program_driver_matches_oracle_bitwise(tmp_path):
    """The stand-alone fvm.f90 program (nx=200, Sod, zero gradient, gravity
    source) over the drop-in compute_update / compute_max_speed: its initial
    state equals the reference's condinit bit for bit, and the run equals the
    oracle's loop over fvm.f90's compute_update bit for bit."""
    tend = 0.05
    prefix = str(tmp_path / "p") + "/"
    os.makedirs(prefix)
    r = run([exe("fvm_fvwb"), repr(tend), prefix])
    assert r.returncode == 0, r.stdout + r.stderr
    with open(prefix + ".bin", "rb") as f:
        nx, nvar, steps = np.frombuffer(f.read(12), dtype=np.int32)
        t = float(np.frombuffer(f.read(8), dtype=np.float64)[0])
        arrs = np.frombuffer(f.read(2 * 3 * 200 * 8), dtype=np.float64)
    assert (nx, nvar) == (200, 3)
    u0, u = (arrs[k * 600:(k + 1) * 600].reshape((3, 200), order="F") for k in range(2))
    g = np.load(os.path.join(ROOT, "tests", "golden", "fvm_update_200.npz"))
    assert np.array_equal(u0, g["u_sod"])
    p = po.make_params_1d(200, bc=po.BC1D_ZERO_GRADIENT)
    uo, to, n = oracle_heun_1d(lambda v: po.fvm_update(p, v, 2), u0, tend, 1.0 / 200, 7.0)
    print(f"steps {steps} vs {n}, t {t} vs {to}, max|du| {np.max(np.abs(u - uo)):.3e}", flush=True)
    assert steps == n and t == to
    assert np.array_equal(u, uo)
    for name in ("value00000.dat", "value99999.dat"):
        lines = open(prefix + name).read().splitlines()
        assert len(lines) == 200 and all(re.fullmatch(r"(?:[ -]\d\.\d{5}E[+-]\d\d ?){4}", ln) for ln in lines)


@pytest.mark.gpu
def test_1d_driver_resident_run_matches_oracle_bitwise(tmp_path):
    """Resident mode runs the configured scheme (MUSCL, Rusanov, SSP-RK2, walls,
    CFL 0.5) to tend in one launch: equal to the oracle's run bit for bit."""
    tend = 0.1
    prefix = str(tmp_path / "r1")
    r = run([exe("benchmark_1d_fvwb"), "resident", repr(tend), prefix])
    assert r.returncode == 0, r.stdout + r.stderr
    steps, t, u0, w_eq, u = read_state_1d(prefix + ".bin")
    p = po.make_params_1d(128, recon=po.RECON_MUSCL_MINMOD, bc=po.BC1D_WALL, integrator=po.RK2_HEUN, cfl=0.5)
    uo, to, no, _ = po.fv1d_advance(p, u0, w_eq, t_end=tend)
    assert steps == no and t == to == tend
    assert np.array_equal(u, uo)
    lines = open(prefix + "_final.dat").read().splitlines()
    assert len(lines) == 128 and all(re.fullmatch(r"(?:[ -]\d\.\d{5}E[+-]\d\d ?){2}", ln) for ln in lines)


def dg1d_golden():
    return np.load(os.path.join(ROOT, "tests", "golden", "dg1d_run_rki_128.npz"))


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["stage", "resident"])
def test_dg1d_driver_reproduces_reference_program(tmp_path, mode):
    """The 1D DG program (n = 3, nx = 128, HLLC, ninit = 8, 'RKi', tend = 0.2)
    with the drop-in residual (stage) or the device-resident loop (resident):
    the same 236 steps as the reference program, every printed t and dt (stage)
    or the last t (resident) equal to the reference's bit for bit, the same
    printed 'error', and recon99999.dat equal to the reference's."""
    g = dg1d_golden()
    prefix = str(tmp_path) + "/"
    r = run([exe("dg1d_fvwb"), mode, "0.2", prefix])
    assert r.returncode == 0, r.stdout + r.stderr
    err = float(re.findall(r"error\s+(\S+)", r.stdout)[-1])
    steps, n, nvar = None, 3, 3
    with open(prefix + ".bin", "rb") as f:
        nx, n, nvar, steps = np.frombuffer(f.read(16), dtype=np.int32)
        t, dt = np.frombuffer(f.read(16), dtype=np.float64)
    print(f"{mode}: steps {steps}, t {t!r}, error {err!r}", flush=True)
    assert (nx, n, nvar) == (128, 3, 3) and steps == 236
    assert t == g["t"][-1] and dt == g["dt"][-1]
    assert err == float(g["error"])
    if mode == "stage":
        rows = re.findall(r"time=\s+(\d+)\s+(\S+)\s+(\S+)", r.stdout)
        np.testing.assert_array_equal([float(b) for _, b, _ in rows], g["t"])
        np.testing.assert_array_equal([float(c) for _, _, c in rows], g["dt"])
    np.testing.assert_array_equal(np.loadtxt(prefix + "recon99999.dat"), g["recon"])


@pytest.mark.gpu
def test_driver_resume_from_checkpoint_equals_one_run(tmp_path):
    """The resident driver saves its device state at t = 0.01 through the
    Fortran binding of fvwb_checkpoint_save; a second run resumes from that
    file to t = 0.02 and ends exactly where one run to 0.02 ends."""
    ck = str(tmp_path / "s.ckpt")
    r = run([exe("benchmark_2d_fvwb_64x48"), "resident", "0.01", str(tmp_path / "a"), ck])
    assert r.returncode == 0, r.stdout + r.stderr
    r = run([exe("benchmark_2d_fvwb_64x48"), "resume", "0.02", str(tmp_path / "b"), ck])
    assert r.returncode == 0, r.stdout + r.stderr
    r = run([exe("benchmark_2d_fvwb_64x48"), "resident", "0.02", str(tmp_path / "c")])
    assert r.returncode == 0, r.stdout + r.stderr
    sb, tb, _, _, ub = read_state(str(tmp_path / "b.bin"))
    sc, tc, _, _, uc = read_state(str(tmp_path / "c.bin"))
    sa, ta, _, _, _ = read_state(str(tmp_path / "a.bin"))
    print(f"checkpoint at step {sa} (t {ta}); resumed {sb} steps to {tb}; one run {sc} to {tc}", flush=True)
    assert 0 < sa < sb and sb == sc and tb == tc
    assert np.array_equal(ub, uc)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["stage", "resident"])
def test_rd_driver_matches_oracle(tmp_path, mode):
    """The residual-distribution program (65x65 nodes, ninit = 8, transmissive
    copies, 'RD1', tend = 0.005) with the drop-ins on the host loop (stage) or
    the device-resident loop (resident): from the driver's own initial state
    the oracle's restatement of evolve takes the same steps to the same t and
    state, bit for bit; the initial state itself equals the oracle's to libm
    rounding."""
    prefix = str(tmp_path / "rd")
    r = run([exe("rd_fvwb"), mode, "0.005", prefix])
    assert r.returncode == 0, r.stdout + r.stderr
    with open(prefix + ".bin", "rb") as f:
        nx, ny, steps, _ = np.frombuffer(f.read(16), dtype=np.int32)
        t = np.frombuffer(f.read(8), dtype=np.float64)[0]
        u0 = np.frombuffer(f.read(8 * 4 * nx * ny), dtype=np.float64).reshape((4, nx, ny), order="F")
        u = np.frombuffer(f.read(8 * 4 * nx * ny), dtype=np.float64).reshape((4, nx, ny), order="F")
    op = po.make_rd_params(int(nx), int(ny), bc=po.RD_BC_TRANSMISSIVE, cfl=0.5)
    np.testing.assert_allclose(u0, po.rd_initial(op, 8, float(np.float32(0.01))), rtol=1e-15, atol=0)
    uo, to, no = po.rd_run(op, np.asfortranarray(u0), 0.005)
    print(f"{mode}: steps {steps} vs {no}, t {t!r} vs {to!r}", flush=True)
    assert (nx, ny) == (65, 65) and steps == no == 162 and t == to
    np.testing.assert_array_equal(u, uo)


def read_rd_bin(prefix):
    with open(prefix + ".bin", "rb") as f:
        nx, ny, steps, _ = np.frombuffer(f.read(16), dtype=np.int32)
        t = np.frombuffer(f.read(8), dtype=np.float64)[0]
        f.read(8 * 4 * nx * ny)
        u = np.frombuffer(f.read(8 * 4 * nx * ny), dtype=np.float64).reshape((4, nx, ny), order="F")
    return int(steps), float(t), u


@pytest.mark.gpu
def test_rd_driver_resume_from_checkpoint_equals_one_run(tmp_path):
    """The resident residual-distribution driver stops after 58 steps of a
    run to t = 0.005 (a step boundary of the uninterrupted run, so no step is
    clipped) and saves its device state and clock through the Fortran binding
    of fvwb_rd_checkpoint_save; a second run loads that file
    (fvwb_rd_checkpoint_load) and continues to t = 0.005, ending with the
    steps, time and state of one run to 0.005, bit for bit."""
    ck = str(tmp_path / "rd.ckpt")
    r = run([exe("rd_fvwb"), "resident", "0.005", str(tmp_path / "a"), ck, "58"])
    assert r.returncode == 0, r.stdout + r.stderr
    r = run([exe("rd_fvwb"), "resume", "0.005", str(tmp_path / "b"), ck])
    assert r.returncode == 0, r.stdout + r.stderr
    r = run([exe("rd_fvwb"), "resident", "0.005", str(tmp_path / "c")])
    assert r.returncode == 0, r.stdout + r.stderr
    sa, ta, _ = read_rd_bin(str(tmp_path / "a"))
    sb, tb, ub = read_rd_bin(str(tmp_path / "b"))
    sc, tc, uc = read_rd_bin(str(tmp_path / "c"))
    print(f"rd checkpoint at step {sa} (t {ta}); resumed to {sb} steps, t {tb}; one run {sc}, t {tc}", flush=True)
    assert sa == 58 and 0 < ta < tb and sb == sc == 162 and tb == tc
    np.testing.assert_array_equal(ub, uc)


DG_DRIVERS = {"benchmark_2d_dg_fvwb": "shipped", "benchmark_2d_dg_fvwb_llf1": "llf1",
              "benchmark_2d_dg_fvwb_hll2": "hll2"}


def read_dg_bin(path):
    with open(path, "rb") as f:
        nx, ny, mx, steps = (int(v) for v in np.frombuffer(f.read(16), dtype=np.int32))
        t = float(np.frombuffer(f.read(8), dtype=np.float64)[0])
        u = np.frombuffer(f.read(8 * 4 * nx * ny * mx * mx), dtype=np.float64).reshape((4, nx, ny, mx, mx), order="F")
    return steps, t, u


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["stage", "resident"])
@pytest.mark.parametrize("name", list(DG_DRIVERS))
def test_dg2d_driver_reproduces_reference_program(tmp_path, name, mode):
    """The reference's 2D DG program over the GPU drop-ins, run to the fixture's
    tend with the reference's host loop (stage) or resident on the device:
    the steps, t and final nodal state equal the reference build's evolve bit
    for bit; the snapshot files it writes into longrotation2_o1/ (derp.dat,
    and SIM00001.dat after step 50) equal the reference's byte for byte; the
    printed compute_error norms equal the reference's to the device
    reduction's rounding."""
    with np.load(os.path.join(ROOT, "tests", "golden", f"dgq_{DG_DRIVERS[name]}.npz")) as z:
        g = {k: z[k] for k in z.files}
    os.makedirs(tmp_path / "longrotation2_o1")
    r = run([exe(name), mode, repr(float(g["tend"])), str(tmp_path / "out")], cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    steps, t, u = read_dg_bin(str(tmp_path / "out.bin"))
    print(f"{name} {mode}: {steps} steps to t {t!r}; reference {len(g['evolve_dt'])}", flush=True)
    assert steps == len(g["evolve_dt"]) and t == float(g["tend"])
    np.testing.assert_array_equal(u, g["evolve_u"])
    for f in g["files"]:
        with open(tmp_path / "longrotation2_o1" / str(f)) as fh:
            assert fh.read() == str(g["file_" + str(f)]), f
    printed = dict(zip([str(k) for k in g["norm_keys"]], g["norms"]))
    for key, val in printed.items():
        m = re.findall(re.escape(key) + r"\s+(\S+)", r.stdout)
        np.testing.assert_allclose(float(m[-1]), val, rtol=1e-13, atol=1e-300, err_msg=key)
    if mode == "stage":
        assert len(re.findall(r"time=", r.stdout)) == steps


@pytest.mark.gpu
def test_1d_driver_snapshot_cadence(tmp_path):
    """make_movie in the 1D driver: simul/SIMnnn.dat after every 200th step
    (benchmark_1d.f90:252-256), each the state of that step in output_file's
    format (x and p - p_eq, '(7(1PE12.5,1X))'); the state comes from the
    oracle's restatement of the same loop, which the drop-in loop equals bit
    for bit."""
    os.makedirs(tmp_path / "simul")
    tend = 1.5
    r = run([exe("benchmark_1d_fvwb_64_movie"), "stage", repr(tend), str(tmp_path / "m")], cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    steps, t, u0, w_eq, u = read_state_1d(str(tmp_path / "m.bin"))
    files = sorted(os.listdir(tmp_path / "simul"))
    print(f"{steps} steps, snapshots {files}", flush=True)
    assert steps >= 400 and files == [f"SIM{k:03d}.dat" for k in range(1, steps // 200 + 1)]
    p = po.make_params_1d(64)
    dx = 1.0 / 64
    x = (np.arange(64) + 0.5) * dx
    uo, to, n = np.asfortranarray(u0.copy()), 0.0, 0
    while n < 200:
        cmax = po.ref1d_max_speed(p, uo)
        dt = float(np.float32(0.8)) * dx / cmax / (2.0 * 1.0 + 1.0)
        d = po.ref1d_update(p, uo, w_eq)
        w1 = uo + dt * d
        d = po.ref1d_update(p, w1, w_eq)
        uo = 0.5 * uo + 0.5 * w1 + 0.5 * dt * d
        n += 1
    gm = p.gamma
    pr = (gm - 1.0) * (uo[2] - 0.5 * uo[0] * (uo[1] / uo[0]) ** 2)
    want = "".join(f"{a:12.5E} {b:12.5E}\n" for a, b in zip(x, pr - w_eq[2]))
    with open(tmp_path / "simul" / "SIM001.dat") as f:
        assert f.read() == want
