"""Benchmark: cell-updates/s of the well-balanced Euler-with-gravity solver on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--drift-steps D]

Prints ONE JSON line.  A step advances every instance by one full SSP-RK3 time
step (three persistent stage launches, dt per instance from the on-device CFL
reduction); inputs are resident in HBM before the timed region, which is
bracketed by HIP events on the solver's stream and a device synchronisation on
both sides.  A cell-update is one cell advanced through one full step.

Workload (`value`): a batch of 512 independent 2D gravity problems, sides drawn
log-uniform in [64, 1024] (seed 20261202), g ~ U[0.5, 2] (phi = g(x+y),
rho = p = exp(-g(x+y))), Gaussian pressure bump of amplitude ~ U[1e-4, 1e-2];
FP32-bulk deviation form, MUSCL-minmod, HLL, reflecting walls, SSP-RK3; launch
shape chosen by the library's autotuner.  Beside it: the plain FP64 kernel on
the same batch (speedup and L1 density error of FP32 vs FP64), the 8192x8192
single problem in both precisions, a pure-equilibrium drift run, the roofline
of the stage kernel; then the P2 modal DG workload (1024^2, 100 steps), the
1D workload (256 cells to t = 0.25, alone and 4096 side by side) and the
residual-distribution workload.

CPU leg (skipped with --no-cpu): the Fortran FP64 restatement of the configured
scheme (oracle/fvwb_cpu_fv2d.f90, flang -O3) timed on one core on a bounded
sample of the batch (`cpu_baseline`) and on all cores at 2048^2; the
reference's own benchmark_2d.f90 (flang -O3 build at 2048^2, oracle/_ref)
timed beside the GPU running the same scheme; and, per workload, the L1 and
L-infinity density error of the GPU results against these FP64 CPU runs and
the FP64 C oracle.
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fvm-source-wb_amd"))

import numpy as np  # noqa: E402

import fvwb  # noqa: E402

HBM_PEAK_GBS = 8000.0
PORT = os.path.join(ROOT, "oracle", "_build", "fvwb_cpu_fv2d")
REF_TIME = os.path.join(ROOT, "oracle", "ref_time.py")
METRIC = "cell-updates/s on 1 MI355X: batched FP32-bulk vs FP64 plain at stated L1 err"
BYTES_PER_CELL_UPDATE = {fvwb.FP32: 128, fvwb.FP64: 256}


class HipEvents:
    """hipEvent timing on an explicit stream (the solver's own)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        for n in ("hipEventCreate", "hipEventRecord", "hipEventSynchronize", "hipEventElapsedTime",
                  "hipEventDestroy", "hipDeviceSynchronize"):
            getattr(self.hip, n).restype = ctypes.c_int
        self.hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def event(self):
        e = ctypes.c_void_p()
        assert self.hip.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.hip.hipEventRecord(e, ctypes.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.hip.hipEventElapsedTime(ctypes.byref(ms), a, b) == 0
        return float(ms.value)

    def sync(self):
        assert self.hip.hipDeviceSynchronize() == 0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def batch_workload(n=512, seed=20261202):
    rng = np.random.default_rng(seed)
    sides = np.rint(np.exp(rng.uniform(np.log(64), np.log(1024), n))).astype(int)
    gs = rng.uniform(0.5, 2.0, n)
    amps = rng.uniform(1e-4, 1e-2, n)
    return sides, gs, amps


def scheme_params(nx, ny, precision, g=1.0):
    return fvwb.make_params(nx, ny, recon=fvwb.RECON_MUSCL_MINMOD, flux=fvwb.FLUX_HLL, bc=fvwb.BC_REFLECT,
                            integrator=fvwb.SSPRK3, precision=precision, gamma=1.4, rho0=1.0, p0=1.0, g=g)


def time_batch(ev, sides, gs, amps, precision, steps, warmup, shape=None, keep=()):
    """Times `steps` steps of the batch; returns (seconds, batch, launch shape,
    {instance: initial state} for the instances in `keep`)."""
    b = fvwb.Batch2D([scheme_params(int(s), int(s), precision, float(g)) for s, g in zip(sides, gs)])
    b.init_gaussian(amps)
    if shape is None:
        shape = b.autotune(3, None)
    else:
        b.set_launch(shape[0], shape[1], 0)
    if warmup:
        b.advance(warmup)
    b.init_gaussian(amps)
    u0 = {k: b.download(k) for k in keep}
    st = b.stream()
    e0, e1 = ev.event(), ev.event()
    ev.sync()
    ev.record(e0, st)
    b.advance_async(steps)
    ev.record(e1, st)
    ms = ev.elapsed_ms(e0, e1)
    ev.sync()
    for k in (0, len(sides) - 1):
        assert b.time(k)[2] == steps
    return ms / 1e3, b, shape, u0


def time_single(ev, n, precision, steps, warmup):
    s = fvwb.Solver2D(scheme_params(n, n, precision))
    s.init_gaussian(1e-3, 0.3, 0.3, 100.0)
    if warmup:
        s.advance(warmup)
    s.init_gaussian(1e-3, 0.3, 0.3, 100.0)
    st = s.stream()
    e0, e1 = ev.event(), ev.event()
    ev.sync()
    ev.record(e0, st)
    s.advance_async(steps)
    ev.record(e1, st)
    ms = ev.elapsed_ms(e0, e1)
    ev.sync()
    assert s.time()[2] == steps
    return ms / 1e3, s


def measured_traffic(cells, shape):
    """HBM bytes per stage launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs (scaled by the copy-kernel calibration), for the workload
    and launch shape they were taken on; None for a shape not profiled."""
    path = os.path.join(ROOT, "profiles", "fv2d_batch512_fp32_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except OSError:
        return None, "no profile"
    key = f"{shape[0]},{shape[1]}"
    if d.get("cells") != cells or key not in d.get("shapes", {}):
        return None, f"no profile for shape {list(shape)} (profiled: {sorted(d.get('shapes', {}))})"
    return d["shapes"][key]["hbm_bytes_per_launch"], f"profiles/fv2d_batch512_fp32_traffic.json shape {key}"


def time_dg(ev, n, steps, warmup, integrator=fvwb.DG_SSPRK3):
    """The DG workload: n x n cells, P2 (6 modes, 3x3 Gauss points), HLL, TVB
    (M = 50), walls, SSP-RK3 (or the reference's SSP(5,4)), g = 1 atmosphere
    plus a 16-term Fourier pressure perturbation of amplitude 1e-4 (seed 0xF00D)."""
    a, mk, nk, th = fvwb.dg_fourier_terms()
    p = fvwb.make_dg_params(n, n, integrator=integrator)
    s = fvwb.SolverDG(p)
    s.init_fourier(a, mk.astype(float), nk.astype(float), th, 1e-4)
    if warmup:
        s.advance(warmup)
        s.init_fourier(a, mk.astype(float), nk.astype(float), th, 1e-4)
    st = s.stream()
    e0, e1 = ev.event(), ev.event()
    ev.sync()
    ev.record(e0, st)
    s.advance_async(steps)
    ev.record(e1, st)
    ms = ev.elapsed_ms(e0, e1)
    ev.sync()
    assert s.time()[2] == steps
    s.close()
    return ms / 1e3


def time_dgq(n, order=3, steps=20):
    """The reference's own 2D DG program on the device (fvwb_dgq_*, FP64,
    tensor basis with mx = my = order): n x n cells on the unit square, 'hll2',
    get_source with the point-mass grad_phi, periodic copies, solver 'RK4' with
    apply_limiter ('ONP') after every stage and the dt of
    2d/benchmark_2d_dg.f90:671; a smooth state rho = 1 + 0.2 sin(2 pi x)
    sin(2 pi y), u = v = 0.1, p = 1 at the Gauss nodes.  The resident loop
    reads its clock between groups of 32 steps, so the host clock times it."""
    p = fvwb.reference_params_dgq(n, order=order, flux_type=fvwb.DGQ_FLUX_HLL2, source=2)
    xq, _ = fvwb.dgq_gl_quadrature(order)
    dx, dy = p.boxlen_x / n, p.boxlen_y / n
    ix = np.arange(n)
    x = (ix + 0.5)[:, None, None, None] * dx + 0.5 * dx * xq[None, None, :, None] + 0.0 * xq[None, None, None, :]
    y = (ix + 0.5)[None, :, None, None] * dy + 0.5 * dy * xq[None, None, None, :] + 0.0 * xq[None, None, :, None]
    x = np.asfortranarray(np.broadcast_to(x, (n, n, order, order)))
    y = np.asfortranarray(np.broadcast_to(y, (n, n, order, order)))
    rho = 1.0 + 0.2 * np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y)
    nodes = np.zeros((4, n, n, order, order), order="F")
    nodes[0] = rho
    nodes[1] = 0.1 * rho
    nodes[2] = 0.1 * rho
    nodes[3] = 1.0 / (p.gamma - 1.0) + 0.5 * rho * 0.02
    with fvwb.SolverDGQ(p) as s:
        du = s.apply_limiter(s.modes_from_nodes(nodes))
        s.upload(du, x, y)
        s.advance(2)                     # code objects loaded outside the timed region
        s.upload(du, x, y)
        t0 = time.perf_counter()
        s.advance(steps)
        sec = time.perf_counter() - t0
        assert s.time()[2] == steps
    return sec


def time_fv1d(ev, ninst, n=256, t_end=0.25):
    """The 1D gravity workload: n cells, MUSCL-minmod, Rusanov, SSP-RK2, walls,
    CFL 0.5, rho = p = exp(-x) plus a pressure pulse 1e-3 exp(-(x-0.5)^2/(2 0.05^2)),
    run to t_end; `ninst` copies run side by side, one workgroup each."""
    ps = [fvwb.make_params_1d(n) for _ in range(ninst)]
    s = fvwb.Solver1D(ps)
    s.init_gaussian(np.full(ninst, 1e-3), 0.5, 0.05)
    s.advance(nsteps=2)                  # loads the kernel's code object outside the timed region
    s.init_gaussian(np.full(ninst, 1e-3), 0.5, 0.05)
    st = s.stream()
    e0, e1 = ev.event(), ev.event()
    ev.sync()
    ev.record(e0, st)
    s.advance(t_end=t_end)
    ev.record(e1, st)
    ms = ev.elapsed_ms(e0, e1)
    t, _, nsteps = s.time(0)
    assert t == t_end
    s.close()
    return ms / 1e3, nsteps


def time_rd(n, steps, eta=1e-2):
    """Residual distribution (benchmark_rd.f90, solver 'RD1'): n x n nodes on
    the unit square, transmissive copies, CFL 0.5, the rest state rho = 1.21
    exp(-1.21 (x+y)) with a pressure bump eta exp(-121 ((x-0.3)^2 + (y-0.3)^2))
    (initial condition 8, built here in numpy).  Each step reads the speed
    maximum back to the host for dt, as the reference's loop does, so the
    time is taken on the host clock around the synchronous advance."""
    gamma = float(np.float32(1.4))
    x = np.linspace(0.0, 1.0, n)[:, None]
    y = np.linspace(0.0, 1.0, n)[None, :]
    r = float(np.float32(1.21))
    u = np.zeros((4, n, n), order="F")
    u[0] = r * np.exp(-r * (x + y))
    u[3] = (np.exp(-r * (x + y)) + eta * np.exp(-121.0 * ((x - 0.3) ** 2 + (y - 0.3) ** 2))) / (gamma - 1.0)
    with fvwb.SolverRD(fvwb.make_rd_params(n, n, gamma=gamma)) as s:
        s.upload(u)
        s.advance(2)                         # code objects loaded outside the timed region
        s.upload(u)
        t0 = time.perf_counter()
        s.advance(steps)
        sec = time.perf_counter() - t0
        assert s.time()[2] == steps
    return sec


def cpu_leg(fn, *a):
    """Runs one part of the CPU leg; a failure there (a missing build, a host
    limit) is recorded in the JSON line instead of ending the run."""
    try:
        return fn(*a)
    except Exception as e:  # noqa: BLE001
        log(f"CPU leg {fn.__name__} failed: {e!r}")
        return {"error": f"{fn.__name__}: {e!r}"}


def host_info():
    """CPU model and the cores this process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": len(os.sched_getaffinity(0)), "cpu_count": os.cpu_count()}


def density_err(a, b):
    """L1 (mean over the unit square, i.e. sum |d| dx dy) and L-infinity of a
    density difference."""
    d = np.abs(np.asarray(a, dtype=np.float64) - np.asarray(b, dtype=np.float64))
    return {"l1": float(d.mean()), "linf": float(d.max())}


def run_port(u, w, steps, g, threads, want_state=True):
    """The Fortran FP64 restatement of the configured scheme
    (oracle/fvwb_cpu_fv2d.f90: MUSCL-minmod, HLL, walls, SSP-RK3, deviation
    form, gamma 1.4, CFL 0.5) on one problem from state u with equilibrium w;
    returns (state after `steps` steps or None, seconds of its time loop)."""
    nx, ny = u.shape[1], u.shape[2]
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        fout = os.path.join(td, "out.bin") if want_state else "-"
        with open(fin, "wb") as f:
            f.write(np.asfortranarray(u, dtype=np.float64).tobytes(order="F"))
            f.write(np.asfortranarray(w, dtype=np.float64).tobytes(order="F"))
        r = subprocess.run([PORT, str(nx), str(ny), str(steps), repr(1.4), repr(0.5), "1.0", "1.0", repr(float(g)),
                            fin, fout, str(threads)], capture_output=True, text=True, check=True)
        sec = float(r.stdout.split()[1])
        out = np.fromfile(fout).reshape((4, nx, ny), order="F") if want_state else None
    return out, sec


def cpu_batch_sample(sides, gs, u0, steps, b32, b64):
    """cpu_baseline: the port on one core over a bounded sample of the batch
    (the first instances in batch order, each run the full `steps` steps from
    the batch's own initial state), and the FP32 and FP64 batch results of
    those instances against it."""
    import pyoracle as po

    cells, sec, errs32, errs64 = 0, 0.0, [], []
    for k, uk in u0.items():
        s, g = int(sides[k]), float(gs[k])
        op = po.make_params(s, s, recon=po.RECON_MUSCL_MINMOD, flux=po.FLUX_HLL, bc=po.BC_REFLECT,
                            integrator=po.SSPRK3, gamma=1.4, rho0=1.0, p0=1.0, g=g)
        ref, t = run_port(uk, po.equilibrium_primitive(op), steps, g, 1)
        cells += s * s
        sec += t
        errs32.append(density_err(b32.download(k)[0], ref[0]))
        errs64.append(density_err(b64.download(k)[0], ref[0]))
    worst = lambda es, key: max(e[key] for e in es)  # noqa: E731
    return {"value": cells * steps / sec, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"oracle/fvwb_cpu_fv2d.f90 (Fortran FP64 restatement of the same scheme, flang -O3 "
                      f"-ffp-contract=off -march=x86-64-v3), one thread, instances {sorted(u0)} of the batch "
                      f"(sides {[int(sides[k]) for k in sorted(u0)]}, {cells} cells) for the full {steps} steps "
                      f"({sec:.1f} s)",
            "host": host_info(),
            "accuracy_vs_port": {"instances": sorted(u0),
                                 "fp32_batch": {"l1_max": worst(errs32, "l1"), "linf_max": worst(errs32, "linf")},
                                 "fp64_batch": {"l1_max": worst(errs64, "l1"), "linf_max": worst(errs64, "linf")}}}


def time_solver(ev, s, u, w, steps):
    """hipEvent time of `steps` steps of a Solver2D from state (u, w)."""
    s.upload(u, w)
    s.advance(2)                              # code objects loaded outside the timed region
    s.upload(u, w)
    st = s.stream()
    e0, e1 = ev.event(), ev.event()
    ev.sync()
    ev.record(e0, st)
    s.advance_async(steps)
    ev.record(e1, st)
    ms = ev.elapsed_ms(e0, e1)
    ev.sync()
    assert s.time()[2] == steps
    return ms / 1e3


def config_2048(ev, steps, threads):
    """The 2048^2 FP64 problem (MUSCL-minmod, HLL, walls, SSP-RK3, exp(-(x+y))
    with a 1e-3 pressure bump): `steps` steps on the GPU in FP64 and in the
    FP32-bulk form, the port on `threads` cores for the same steps (and on one
    core for 2 steps), and the GPU densities against the port's."""
    import pyoracle as po

    n = 2048
    op = po.make_params(n, n, recon=po.RECON_MUSCL_MINMOD, flux=po.FLUX_HLL, bc=po.BC_REFLECT,
                        integrator=po.SSPRK3, gamma=1.4, rho0=1.0, p0=1.0, g=1.0)
    u, w = po.perturbed_state(op, amp_bump=1e-3, amp_noise=0.0)
    out = {"workload": f"{n}x{n}, exp(-(x+y)) atmosphere, radial 1e-3 pressure bump at (0.3, 0.3); MUSCL-minmod, "
                       f"HLL, walls, SSP-RK3, CFL 0.5", "steps": steps}
    dens = {}
    for name, prec in (("gpu_fp64", fvwb.FP64), ("gpu_fp32", fvwb.FP32)):
        with fvwb.Solver2D(scheme_params(n, n, prec)) as s:
            sec = time_solver(ev, s, u, w, steps)
            dens[name] = s.download()[0]
        out[name] = {"value": n * n * steps / sec, "unit": "cell-updates/s", "ms_per_step": sec * 1e3 / steps}
    ref, sec_all = run_port(u, w, steps, 1.0, threads)
    _, sec_one = run_port(u, w, 2, 1.0, 1, want_state=False)
    out["cpu_port"] = {"value": n * n * steps / sec_all, "unit": "cell-updates/s", "cores": threads,
                       "ms_per_step": sec_all * 1e3 / steps, "kind": "port"}
    out["cpu_port_1core"] = {"value": n * n * 2 / sec_one, "unit": "cell-updates/s", "cores": 1,
                             "sample": f"2 steps ({sec_one:.1f} s)", "kind": "port"}
    out["speedup_gpu_fp64_vs_cpu_1core"] = out["gpu_fp64"]["value"] / out["cpu_port_1core"]["value"]
    out["speedup_gpu_fp64_vs_cpu_all_cores"] = out["gpu_fp64"]["value"] / out["cpu_port"]["value"]
    out["fp64_vs_port"] = density_err(dens["gpu_fp64"], ref[0])
    out["fp32_vs_port"] = density_err(dens["gpu_fp32"], ref[0])
    return out


def reference_leg(ev, gpu_steps):
    """The reference's own benchmark_2d.f90 (evolve: first-order Rusanov,
    frozen ring, Heun, exp(-(x+y)) equilibrium) built with flang -O3 at
    2048^2, run for 2 steps on one core in a child process, beside the GPU
    drop-in of the same scheme (FP64) timed over `gpu_steps` steps."""
    import pyoracle as po

    n = 2048
    op = po.make_params(n, n, rho0=1.0)
    u, w = po.perturbed_state(op, amp_bump=1e-3, amp_noise=0.0)
    c = po.ref2d_max_speed(op, u)
    tend = 1.5 * (0.5 * (1.0 / n) / c * op.cfl)        # two steps of dt = 0.5*dx/cmax*cfl
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.npz"), os.path.join(td, "out.npz")
        np.savez(fin, u=u, w_eq=w, tend=tend)
        r = subprocess.run([sys.executable, REF_TIME, fin, fout], capture_output=True, text=True, check=True)
        steps = sum(1 for line in r.stdout.splitlines() if "time=" in line)
        with np.load(fout) as z:
            uref, sec = z["u"], float(z["seconds"])
    with fvwb.Solver2D(fvwb.reference_params(n, n, 1)) as s:
        s.upload(u, w)
        s.advance(1 << 30, t_end=tend)
        ng = s.time()[2]
        diff = float(np.max(np.abs(s.download() - uref)))
        gsec = time_solver(ev, s, u, w, gpu_steps)
    cpu = n * n * steps / sec
    gpu = n * n * gpu_steps / gsec
    return {"value": cpu, "unit": "cell-updates/s", "cores": 1, "kind": "reference",
            "sample": f"oracle/_ref/libref_bench2d_2048.so: benchmark_2d.f90 evolve (Heun over "
                      f"compute_update_exact, first-order Rusanov, frozen ring) at {n}x{n}, flang -O3 "
                      f"-march=x86-64-v3, its whole-array dump of dudt left out; {steps} steps ({sec:.1f} s)",
            "gpu_same_scheme": {"value": gpu, "unit": "cell-updates/s", "steps": gpu_steps,
                                "ms_per_step": gsec * 1e3 / gpu_steps, "precision": "fp64",
                                "speedup_vs_reference_cpu": gpu / cpu},
            "max_abs_state_diff_after_run": diff, "gpu_steps_in_check": ng}


def dg_accuracy(n, steps):
    """The DG workload's FP32 GPU run against the FP64 C oracle for `steps`
    steps from the GPU's own initial modes: density error at the 3x3 Gauss
    points, L1 weighted as the reference's compute_error weights it
    (2d/benchmark_2d_dg.f90:45-59) and L-infinity."""
    import pyoracle as po

    a, mk, nk, th = fvwb.dg_fourier_terms()
    with fvwb.SolverDG(fvwb.make_dg_params(n, n)) as s:
        s.init_fourier(a, mk.astype(float), nk.astype(float), th, 1e-4)
        d0 = s.download()
        s.advance(steps)
        dg = s.download()
    op = po.make_dg_params(n, n)
    t0 = time.perf_counter()
    dref, _, k, _ = po.dg2d_advance(op, d0, max_steps=steps)
    sec = time.perf_counter() - t0
    assert k == steps
    diff = np.abs(po.dg2d_nodal(op, dg)[0] - po.dg2d_nodal(op, dref)[0])
    _, wq = po.dg_gl_quadrature(3)
    wts = np.outer(wq, wq).ravel(order="F")              # point a + 3 b carries w_a w_b
    l1 = float(np.sum(diff * wts[None, None, :]) * 0.25 / (n * n))
    return {"steps": steps, "l1": l1, "linf": float(diff.max()), "oracle_seconds": sec,
            "oracle": "oracle/fvwb_oracle_dg2d.c FP64, one thread"}


def fv1d_accuracy():
    """The 1D workload (256 cells to t = 0.25) on the GPU against the FP64 C oracle."""
    import pyoracle as po

    op = po.make_params_1d(256, recon=po.RECON_MUSCL_MINMOD, bc=po.BC1D_WALL, integrator=po.RK2_HEUN,
                           nequilibrium=2, gamma=1.4, cfl=0.5, g=1.0)
    u, w = po.perturbed_state_1d(op, amp_bump=1e-3, x0=0.5, width=0.05)
    uo, to, no, _ = po.fv1d_advance(op, u, w, t_end=0.25)
    with fvwb.Solver1D(fvwb.make_params_1d(256)) as s:
        s.upload(u, w)
        s.advance(t_end=0.25)
        t, _, n = s.time()
        ug = s.download()
    e = density_err(ug[0], uo[0])
    e.update({"steps": n, "t": t, "bitwise_equal": bool(np.array_equal(ug, uo) and n == no and t == to)})
    return e


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU leg (CPU baseline, reference timing and the accuracy checks against them)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the all-core CPU run (0: the cores this process may use, at most 16)")
    ap.add_argument("--drift-steps", type=int, default=10000)
    ap.add_argument("--shape", default=None,
                    help="ROWS,CPL: fixed launch shape for the batch instead of the autotuner")
    ap.add_argument("--skip-single", action="store_true", help="skip the 8192^2 single-problem runs")
    ap.add_argument("--fp32-only", action="store_true", help="time the FP32 batch only (no JSON line)")
    ap.add_argument("--dg-size", type=int, default=1024, help="cells per side of the DG run (0: skip)")
    ap.add_argument("--fv1d-batch", type=int, default=4096, help="1D instances in the batched 1D run (0: skip)")
    ap.add_argument("--dg-only", action="store_true", help="time the DG workload only (no JSON line)")
    ap.add_argument("--rd-size", type=int, default=2049, help="nodes per side of the residual-distribution run (0: skip)")
    ap.add_argument("--rd-only", action="store_true", help="time the residual-distribution workload only (no JSON line)")
    ap.add_argument("--dgq-size", type=int, default=512,
                    help="cells per side of the reference DG program's run (0: skip)")
    ap.add_argument("--dgq-only", action="store_true", help="time the reference DG program only (no JSON line)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    ev = HipEvents()
    K, W = args.steps, args.warmup
    if args.rd_only:       # profiling runs of the residual-distribution kernels alone
        secr = time_rd(args.rd_size, K)
        log(f"rd-only {args.rd_size}^2: {args.rd_size ** 2 * K / secr:.4e} node-updates/s, "
            f"{secr * 1e3 / K:.3f} ms/step")
        return
    if args.dgq_only:      # profiling runs of the reference DG program's kernels alone
        secq = time_dgq(args.dgq_size, steps=K)
        log(f"dgq-only {args.dgq_size}^2: {args.dgq_size ** 2 * K / secq:.4e} cell-updates/s, "
            f"{secq * 1e3 / K:.3f} ms/step")
        return
    if args.dg_only:       # profiling runs of the DG kernels alone
        sec_dg = time_dg(ev, args.dg_size, K, W)
        log(f"dg-only {args.dg_size}^2: {args.dg_size ** 2 * K / sec_dg:.4e} cell-updates/s, "
            f"{sec_dg * 1e3 / K:.3f} ms/step")
        return

    # ---- headline: the 512-instance batch -------------------------------------------
    sides, gs, amps = batch_workload()
    cells = int(np.sum(sides.astype(np.int64) ** 2))
    fixed = tuple(int(v) for v in args.shape.split(",")) if args.shape else None
    # the CPU sample: the first instances in batch order up to 1.5e5 cells
    sample, acc = [], 0
    while not args.fp32_only and not args.no_cpu and acc < 150000:
        sample.append(len(sample))
        acc += int(sides[len(sample) - 1]) ** 2
    sec32, b32, shape32, _ = time_batch(ev, sides, gs, amps, fvwb.FP32, K, W, fixed)
    rate32 = cells * K / sec32
    log(f"batch fp32: {len(sides)} instances, {cells} cells, shape {shape32}, {K} steps in "
        f"{sec32 * 1e3:.2f} ms -> {rate32:.4e} cell-updates/s")
    if args.fp32_only:     # profiling runs of the FP32 batch kernel alone
        log(f"fp32-only: {rate32:.4e} cell-updates/s, shape {shape32}")
        b32.close()
        return
    sec64, b64, shape64, u0_sample = time_batch(ev, sides, gs, amps, fvwb.FP64, K, W, fixed, keep=sample)
    rate64 = cells * K / sec64
    log(f"batch fp64: shape {shape64}, {K} steps in {sec64 * 1e3:.2f} ms -> {rate64:.4e} cell-updates/s")
    l1_max, linf_max = 0.0, 0.0
    for k, n in enumerate(sides):
        d32 = b32.download_deviation(k)[0]
        d64 = b64.download_deviation(k)[0]
        diff = np.abs(d32 - d64)
        l1_max = max(l1_max, float(diff.sum()) / (n * n))
        linf_max = max(linf_max, float(diff.max()))
    log(f"fp32 vs fp64 density: max L1 {l1_max:.3e}, max Linf {linf_max:.3e}")
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        cpu = cpu_leg(cpu_batch_sample, sides, gs, u0_sample, K, b32, b64)
        if "error" not in cpu:
            log(f"cpu baseline (port, 1 core): {cpu['value']:.4e} cell-updates/s; {cpu['sample']}; "
                f"accuracy {cpu['accuracy_vs_port']}")
    b32.close()
    b64.close()

    # ---- the single 8192^2 problem ---------------------------------------------------------
    n1 = 8192
    single = None
    if not args.skip_single:
        s32sec, s32 = time_single(ev, n1, fvwb.FP32, K, W)
        s64sec, s64 = time_single(ev, n1, fvwb.FP64, K, W)
        single32, single64 = n1 * n1 * K / s32sec, n1 * n1 * K / s64sec
        single = {"fp32": single32, "fp64": single64, "speedup": single32 / single64, "steps": K,
                  "fp32_vs_fp64": density_err(s32.download_deviation()[0], s64.download_deviation()[0])}
        log(f"single {n1}^2: fp32 {single32:.4e}, fp64 {single64:.4e} cell-updates/s; "
            f"fp32 vs fp64 density {single['fp32_vs_fp64']}")
        s32.close()
        s64.close()

    # ---- hydrostatic drift: resting atmosphere, FP32 bulk, 8192^2 --------------------------
    drift = None
    if args.drift_steps > 0:
        s = fvwb.Solver2D(scheme_params(n1, n1, fvwb.FP32))
        s.init_gaussian(0.0, 0.3, 0.3, 100.0)
        t0 = time.perf_counter()
        s.advance(args.drift_steps)
        dev = s.download_deviation()
        drift = {"steps": args.drift_steps, "size": n1, "max_abs_drho": float(np.abs(dev[0]).max()),
                 "max_abs_dE": float(np.abs(dev[3]).max()), "seconds": time.perf_counter() - t0}
        s.close()
        del dev
        log(f"drift after {args.drift_steps} steps: {drift}")

    # ---- roofline of the fused stage kernel -----------------------------------------------
    # algorithmic bytes per cell-update: 8 state copies x 4 variables x 4 B (FP32)
    # over the three SSP-RK3 stages; a launch moves on average cells*128/3 B.
    # The timed region holds exactly 3K stage launches back to back, so the
    # average launch duration is sec32 / (3K).
    launch_us = sec32 / (3 * K) * 1e6
    achieved = cells * 128.0 / 3 / (launch_us * 1e-6) / 1e9
    copy_gbs = fvwb.copy_bandwidth(1 << 30, 10)
    log(f"device copy bandwidth {copy_gbs:.1f} GB/s")
    traffic, traffic_note = measured_traffic(cells, shape32)

    # ---- CPU leg: 2048^2 beside the Fortran port, and the reference's own program ------------
    c2048, ref = None, None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        c2048 = cpu_leg(config_2048, ev, 200, threads)
        if "error" not in c2048:
            log(f"2048^2 x 200: gpu fp64 {c2048['gpu_fp64']['value']:.4e}, fp32 {c2048['gpu_fp32']['value']:.4e}, "
                f"cpu port {threads} cores {c2048['cpu_port']['value']:.4e}, 1 core "
                f"{c2048['cpu_port_1core']['value']:.4e} cell-updates/s; fp64 vs port {c2048['fp64_vs_port']}, "
                f"fp32 vs port {c2048['fp32_vs_port']}")
        ref = cpu_leg(reference_leg, ev, 200)
        if "error" not in ref:
            log(f"reference benchmark_2d.f90 at 2048^2: {ref['value']:.4e} cell-updates/s on 1 core; GPU same "
                f"scheme {ref['gpu_same_scheme']['value']:.4e}; max |u_gpu - u_ref| "
                f"{ref['max_abs_state_diff_after_run']:.2e}")

    # ---- DG workload (P2 modal DG, projections on MFMA) ----------------------------------
    dg = None
    if args.dg_size > 0:
        ndg, kdg = args.dg_size, 100
        sec_dg = time_dg(ev, ndg, kdg, 3)
        cu_dg = ndg * ndg * kdg / sec_dg
        # issued MFMA work: 56 v_mfma_f32_16x16x4f32 (2048 flops each) per row of a 16-column
        # strip (16 forward, 24 backward volume, 16 backward face), 14 owned cells per row
        mfma_tflops = cu_dg * 3 * 56 * 2048 / 14 / 1e12
        dg = {"workload": f"{ndg}x{ndg} P2 modal DG, 3x3 Gauss, HLL, TVB M=50, walls, SSP-RK3, FP32 state and tables",
              "value": cu_dg, "unit": "cell-updates/s", "steps": kdg, "ms_per_step": sec_dg * 1e3 / kdg,
              "bytes_per_cell_update": 768, "hbm_frac": cu_dg * 768 / 8e12,
              "mfma_tflops_issued": mfma_tflops, "mfma_frac_of_fp32_peak": mfma_tflops / 157.3}
        sec54 = time_dg(ev, ndg, kdg, 3, fvwb.DG_SSPRK54)
        dg["ssp54"] = {"value": ndg * ndg * kdg / sec54, "ms_per_step": sec54 * 1e3 / kdg,
                       "note": "the reference's SSP(5,4) (2d/benchmark_2d_dg.f90:683-710), 5 stages per step"}
        log(f"dg {ndg}^2: {cu_dg:.4e} cell-updates/s, {sec_dg * 1e3 / kdg:.3f} ms/step; "
            f"SSP(5,4) {sec54 * 1e3 / kdg:.3f} ms/step")
        if not args.no_cpu:
            dg["fp32_vs_fp64_oracle"] = cpu_leg(dg_accuracy, ndg, 2)
            log(f"dg {ndg}^2 fp32 vs the FP64 oracle: {dg['fp32_vs_fp64_oracle']}")

    # ---- the reference's own 2D DG program (FP64, tensor basis) ----------------------------
    dgq = None
    if args.dgq_size > 0:
        nq, kq = args.dgq_size, 20
        secq = time_dgq(nq, 3, kq)
        dgq = {"workload": f"{nq}x{nq}, mx = my = 3, 'hll2', point-mass gravity source, periodic, 'RK4' with "
                           f"'ONP' after every stage, FP64 (the reference's 2d/benchmark_2d_dg.f90 program)",
               "steps": kq, "ms_per_step": secq * 1e3 / kq, "value": nq * nq * kq / secq, "unit": "cell-updates/s"}
        log(f"dgq {nq}^2: {dgq['value']:.4e} cell-updates/s, {dgq['ms_per_step']:.3f} ms/step")

    # ---- 1D workload (256 cells to t = 0.25, whole run in one launch) ---------------------
    rd = None
    if args.rd_size > 0:
        nrd, krd = args.rd_size, 50
        secr = time_rd(nrd, krd)
        # per step: two updates reading u, diff (64 B) and writing dudt (32 B),
        # the two combinations (128 B, 96 B) and the speed pass (32 B) per node
        rd = {"workload": f"{nrd}x{nrd} nodes, residual distribution 'RD1', transmissive, CFL 0.5, FP64",
              "steps": krd, "ms_per_step": secr * 1e3 / krd, "value": nrd * nrd * krd / secr,
              "unit": "node-updates/s", "bytes_per_node_update": 448,
              "hbm_frac": nrd * nrd * krd * 448 / secr / 8e12}
        log(f"rd {nrd}^2: {rd['value']:.4e} node-updates/s, {rd['ms_per_step']:.3f} ms/step")
    fv1d = None
    if args.fv1d_batch > 0:
        sec1, n1s = time_fv1d(ev, 1)
        secb, nbs = time_fv1d(ev, args.fv1d_batch)
        fv1d = {"workload": "256 cells, MUSCL-minmod, Rusanov, SSP-RK2, walls, CFL 0.5, t_end 0.25, FP64",
                "single_run_ms": sec1 * 1e3, "steps": n1s,
                "batch_instances": args.fv1d_batch, "batch_run_ms": secb * 1e3,
                "batch_value": args.fv1d_batch * 256 * nbs / secb, "unit": "cell-updates/s"}
        log(f"fv1d: one run {sec1 * 1e3:.3f} ms ({n1s} steps); {args.fv1d_batch} runs {secb * 1e3:.3f} ms")
        if not args.no_cpu:
            fv1d["vs_fp64_oracle"] = cpu_leg(fv1d_accuracy)
            log(f"fv1d vs the FP64 oracle: {fv1d['vs_fp64_oracle']}")

    out = {
        "metric": METRIC,
        "value": rate32,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": sec32 * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batch of 512 independent 2D Euler+gravity problems, sides log-uniform "
                               "64..1024 (seed 20261202), g~U[0.5,2], bump amplitude~U[1e-4,1e-2]; FP32-bulk "
                               "deviation, MUSCL-minmod, HLL, reflecting walls, SSP-RK3, autotuned launch",
                   "instances": len(sides), "cells": cells, "launch_shape": list(shape32),
                   "parallelism": "single GPU, one persistent launch per RK stage, hipGraph replay"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_note,
                     "kernel": "fv2d_batch_stage_kernel<float,MUSCL,HLL,REFLECT,DEV>",
                     "avg_launch_us": launch_us, "algorithmic_bytes_per_launch": cells * 128.0 / 3,
                     "bytes_per_cell_update": 128, "copy_gbs": copy_gbs,
                     "frac_of_copy": achieved / copy_gbs},
        "cpu_baseline": cpu,
        "fp64_plain": {"value": rate64, "workload": "same batch", "bytes_per_cell_update": 256,
                       "roofline_frac": rate64 * 256 / 8e12},
        "speedup_fp32_vs_fp64": rate32 / rate64,
        "l1_density_err_fp32_vs_fp64": l1_max,
        "linf_density_err_fp32_vs_fp64": linf_max,
        "single_8192": single,
        "hydrostatic_drift": drift,
        "config_2048": c2048,
        "cpu_reference": ref,
        "dg_p2": dg,
        "dgq": dgq,
        "fv1d": fv1d,
        "rd": rd,
    }
    # density error per workload, L1 = sum |rho - rho_ref| dx dy over the unit
    # square (Gauss-weighted per cell for DG), against the FP64 reference named
    acc = {"batch512_fp32_vs_fp64_gpu": {"l1_max": l1_max, "linf_max": linf_max}}
    if cpu and "error" not in cpu:
        acc["batch512_sample_vs_cpu_port"] = cpu["accuracy_vs_port"]
    if single:
        acc["single8192_fp32_vs_fp64_gpu"] = single["fp32_vs_fp64"]
    if c2048 and "error" not in c2048:
        acc["grid2048_fp64_gpu_vs_cpu_port"] = c2048["fp64_vs_port"]
        acc["grid2048_fp32_gpu_vs_cpu_port"] = c2048["fp32_vs_port"]
    if dg and "fp32_vs_fp64_oracle" in dg:
        acc["dg1024_fp32_gpu_vs_fp64_oracle"] = dg["fp32_vs_fp64_oracle"]
    if fv1d and "vs_fp64_oracle" in fv1d:
        acc["fv1d256_gpu_vs_fp64_oracle"] = fv1d["vs_fp64_oracle"]
    out["accuracy"] = acc
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
