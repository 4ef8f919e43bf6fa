"""The timed CPU legs against the C oracle: the Fortran FP64 restatement of the
configured scheme (oracle/fvwb_cpu_fv2d.f90) equals orc_fv2d_advance bit for
bit on one thread and on several, and the -O3 timing build of the reference's
own benchmark_2d.f90 at 2048^2 (oracle/_ref/libref_bench2d_2048.so) runs the
same Heun step as the restatement of its evolve (orc_ref2d_evolve)."""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PORT = os.path.join(ROOT, "oracle", "_build", "fvwb_cpu_fv2d")
REF_TIME = os.path.join(ROOT, "oracle", "ref_time.py")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref_bench2d_2048.so")


def run_port(u, w, steps, g, threads, gamma=1.4, cfl=0.5):
    nx, ny = u.shape[1], u.shape[2]
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.asfortranarray(u).tobytes(order="F"))
            f.write(np.asfortranarray(w).tobytes(order="F"))
        r = subprocess.run([PORT, str(nx), str(ny), str(steps), repr(gamma), repr(cfl), "1.0", "1.0", repr(g),
                            fin, fout, str(threads)], capture_output=True, text=True, check=True)
        fields = r.stdout.split()
        return np.fromfile(fout).reshape((4, nx, ny), order="F"), float(fields[5]), float(fields[7])


needs_port = pytest.mark.skipif(not os.path.exists(PORT), reason="oracle/_build/fvwb_cpu_fv2d not built")


@needs_port
@pytest.mark.parametrize("nx,ny,g,threads", [(64, 48, 1.0, 1), (40, 72, 1.7, 4), (130, 9, 0.6, 3)])
def test_port_equals_c_oracle_bitwise(nx, ny, g, threads):
    p = po.make_params(nx, ny, recon=po.RECON_MUSCL_MINMOD, flux=po.FLUX_HLL, bc=po.BC_REFLECT,
                       integrator=po.SSPRK3, gamma=1.4, rho0=1.0, p0=1.0, g=g)
    u, w = po.perturbed_state(p, amp_bump=3e-2, amp_noise=1e-3, x0=0.6, y0=0.2, salt=5)
    out, t, dt = run_port(u, w, 9, g, threads)
    uo, to, no, dto = po.fv2d_advance(p, u, w, max_steps=9)
    assert no == 9 and t == to and dt == dto
    np.testing.assert_array_equal(out, uo)


@needs_port
def test_port_keeps_rest_to_rounding():
    """At rest the residual vanishes; the full-state SSP-RK3 combinations
    (3/4 u + 1/4 u1, 1/3 u + 2/3 u2) then round u by at most an ulp or two per
    step, as in the C oracle."""
    p = po.make_params(50, 50, recon=po.RECON_MUSCL_MINMOD, flux=po.FLUX_HLL, bc=po.BC_REFLECT,
                       integrator=po.SSPRK3, gamma=1.4, rho0=1.0, p0=1.0, g=1.3)
    w = po.equilibrium_primitive(p)
    u = po.conservative(1.4, w)
    out, _, _ = run_port(u, w, 20, 1.3, 2)
    uo, _, _, _ = po.fv2d_advance(p, u, w, max_steps=20)
    np.testing.assert_array_equal(out, uo)
    assert np.max(np.abs(out - u)) <= 20 * 4 * np.finfo(float).eps * np.max(np.abs(u))


@needs_port
def test_port_rejects_bad_arguments():
    r = subprocess.run([PORT, "3", "10", "1", "1.4", "0.5", "1", "1", "1", "-", "-"], capture_output=True, text=True)
    assert r.returncode != 0 and "at least 4" in r.stdout
    r = subprocess.run([PORT, "10"], capture_output=True, text=True)
    assert r.returncode != 0 and "usage" in r.stdout


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="oracle/_ref/libref_bench2d_2048.so not built")
def test_reference_timing_build_equals_its_restatement():
    """One Heun step of the reference's evolve at 2048^2 against orc_ref2d_evolve
    (nequilibrium 1): the timing build is the reference's program at -O3, which
    may contract products into FMAs, so the states agree to a few ulp of the
    O(1) fluxes times dt/dx."""
    n = 2048
    p = po.make_params(n, n, rho0=1.0)
    u, w = po.perturbed_state(p, amp_bump=1e-3, amp_noise=0.0)
    c = po.ref2d_max_speed(p, u)
    tend = 0.5 * (0.5 * (1.0 / n) / c * p.cfl)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.npz"), os.path.join(td, "out.npz")
        np.savez(fin, u=u, w_eq=w, tend=tend)
        r = subprocess.run([sys.executable, REF_TIME, fin, fout], capture_output=True, text=True, check=True)
        assert sum(1 for line in r.stdout.splitlines() if "time=" in line) == 1
        with np.load(fout) as z:
            uref = z["u"]
    uo, to, no = po.ref2d_evolve(p, 1, u, w, tend)
    assert no == 1
    assert np.max(np.abs(uref - uo)) <= 1e-13
