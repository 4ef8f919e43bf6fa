"""Benchmark: cell-updates/s of the well-balanced Euler-with-gravity solver on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Prints ONE JSON line.  A step is one full SSP-RK3 time step (three fused stage
kernels, dt from the on-device CFL reduction) over the whole workload; inputs are
resident in HBM before the timed region, which is bracketed by HIP events on the
solver's stream and a device synchronisation on both sides.

Workload (FP32-bulk deviation form, MUSCL-minmod, HLL, reflecting walls,
SSP-RK3): the 8192x8192 gravity problem (rho = p = exp(-(x+y)), 1e-3 Gaussian
pressure bump).  Beside it, the plain FP64 kernel on the same grid and on
2048x2048 (the FP64 drop-in), the FP32-vs-FP64 L1 density error, the roofline of
the stage kernel and the FP64 CPU oracle timed on a bounded sample.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fvm-source-wb_amd"))

import numpy as np  # noqa: E402

import fvwb  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "cell-updates/s on 1 MI355X: batched FP32-bulk vs FP64 plain at stated L1 err"


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


def solver_params(n, precision):
    return fvwb.make_params(n, n, recon=fvwb.RECON_MUSCL_MINMOD, flux=fvwb.FLUX_HLL, bc=fvwb.BC_REFLECT,
                            integrator=fvwb.SSPRK3, precision=precision, gamma=1.4, rho0=1.0, p0=1.0, g=1.0)


def time_solver(ev, n, precision, steps, warmup):
    """Returns (seconds for `steps` steps, solver) with the solver left open."""
    s = fvwb.Solver2D(solver_params(n, precision))
    s.init_gaussian(1e-3, 0.3, 0.3, 100.0)
    if warmup:
        s.advance(warmup)
    s.init_gaussian(1e-3, 0.3, 0.3, 100.0)
    stream = s.stream()
    e0, e1 = ev.event(), ev.event()
    ev.sync()
    ev.record(e0, stream)
    s.advance_async(steps)
    ev.record(e1, stream)
    ms = ev.elapsed_ms(e0, e1)
    ev.sync()
    t, dt, done = s.time()
    assert done == steps, (done, steps)
    return ms / 1e3, s


def cpu_baseline():
    """FP64 CPU oracle (plain C port, one core) on a bounded sample: the 2048x2048
    grid of the FP64 configuration, 2 SSP-RK3 steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    n, steps = 2048, 2
    p = po.make_params(n, n, recon=po.RECON_MUSCL_MINMOD, flux=po.FLUX_HLL, bc=po.BC_REFLECT,
                       integrator=po.SSPRK3, gamma=1.4, rho0=1.0)
    u, w = po.perturbed_state(p, amp_bump=1e-3, amp_noise=0.0)
    t0 = time.perf_counter()
    po.fv2d_advance(p, u, w, max_steps=steps)
    sec = time.perf_counter() - t0
    return {"value": n * n * steps / sec, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"oracle/fvwb_oracle_fv2d.c FP64 MUSCL-minmod+HLL+SSP-RK3, {n}x{n} grid, "
                      f"{steps} steps ({sec:.1f} s), 1 thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=8192)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    ev = HipEvents()
    n, K, W = args.size, args.steps, args.warmup
    cells = n * n

    sec32, s32 = time_solver(ev, n, fvwb.FP32, K, W)
    rate32 = cells * K / sec32
    print(f"[bench] fp32 bulk {n}^2: {K} steps in {sec32 * 1e3:.2f} ms -> {rate32:.4e} cell-updates/s",
          file=sys.stderr, flush=True)
    sec64, s64 = time_solver(ev, n, fvwb.FP64, K, W)
    rate64 = cells * K / sec64
    print(f"[bench] fp64 plain {n}^2: {K} steps in {sec64 * 1e3:.2f} ms -> {rate64:.4e} cell-updates/s",
          file=sys.stderr, flush=True)
    d32 = s32.download_deviation()[0]
    d64 = s64.download_deviation()[0]
    l1 = float(np.abs(d32 - d64).sum()) / cells
    linf = float(np.abs(d32 - d64).max())
    s32.close()
    s64.close()
    del d32, d64

    n2 = 2048
    sec2, s2 = time_solver(ev, n2, fvwb.FP64, K, W)
    rate2 = n2 * n2 * K / sec2
    s2.close()
    print(f"[bench] fp64 plain {n2}^2: {rate2:.4e} cell-updates/s", file=sys.stderr, flush=True)

    # roofline of the fused stage kernel: algorithmic bytes per cell-update
    # 8 state copies x 4 variables x 4 B (FP32) over the three SSP-RK3 stages;
    # a launch moves on average cells*128/3 B in ms_per_step/3.
    bytes_per_step = cells * 128.0
    achieved = bytes_per_step / sec32 * K / 1e9
    copy_gbs = fvwb.copy_bandwidth(1 << 30, 10)
    print(f"[bench] device copy bandwidth {copy_gbs:.1f} GB/s", file=sys.stderr, flush=True)
    cpu = None if args.no_cpu else cpu_baseline()

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
        "config": {"workload": f"2D Euler+gravity {n}x{n} FP32-bulk deviation, MUSCL-minmod, HLL, "
                               "reflecting walls, SSP-RK3, 1e-3 Gaussian pressure bump",
                   "cells": cells, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "fv2d_stage_kernel<float,MUSCL,HLL,REFLECT,DEV>",
                     "bytes_per_cell_update": 128, "copy_gbs": copy_gbs,
                     "frac_of_copy": achieved / copy_gbs},
        "cpu_baseline": cpu,
        "fp64_plain": {"value": rate64, "size": n, "bytes_per_cell_update": 256,
                       "roofline_frac": rate64 * 256 / 8e12,
                       "value_2048": rate2},
        "speedup_fp32_vs_fp64": rate32 / rate64,
        "l1_density_err_fp32_vs_fp64": l1,
        "linf_density_err_fp32_vs_fp64": linf,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
