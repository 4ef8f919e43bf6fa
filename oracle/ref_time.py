"""Times the reference's own 2D program on the CPU (TEST INFRASTRUCTURE: the
timed reference leg of the CPU baseline, never used by the product).

Runs evolve (benchmark_2d.f90:221-260: compute_max_speed, two calls of
compute_update_exact per Heun step) from oracle/_ref/libref_bench2d_2048.so,
the flang -O3 build of benchmark_2d.f90 at 2048 x 2048 cells with the
exp(-(x+y)) atmosphere (oracle/Makefile), in this child process: the library
holds its 4 GB of work arrays as static storage, and its whole-array
expressions still take a few GB of stack, so evolve runs on a thread with an
8 GiB stack.  The program prints one `time=` line per step on stdout, which
the caller reads to count the steps.

Usage: python ref_time.py in.npz out.npz
  in.npz : u, w_eq (4, 2048, 2048) float64, tend
  out.npz: u after the run, seconds (wall clock of evolve)
"""
import ctypes
import os
import sys
import threading
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_ref", "libref_bench2d_2048.so")
N = 2048


def main(fin, fout):
    with np.load(fin) as z:
        u = np.asfortranarray(z["u"]).copy(order="F")
        w = np.asfortranarray(z["w_eq"])
        tend = float(z["tend"])
    if u.shape != (4, N, N) or w.shape != (4, N, N):
        raise ValueError(f"the reference build is compiled for (4, {N}, {N}); got {u.shape}")
    lib = ctypes.CDLL(LIB)
    ctypes.c_double.in_dll(lib, "_QMparameters_2dEtend").value = tend
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    res = {}

    def run():
        t0 = time.perf_counter()
        lib.evolve_(dp(u), dp(w))
        res["seconds"] = time.perf_counter() - t0

    threading.stack_size(8 << 30)
    th = threading.Thread(target=run)
    th.start()
    th.join()
    if "seconds" not in res:
        raise RuntimeError("the reference's evolve did not return")
    np.savez(fout, u=u, seconds=res["seconds"])


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2])
