"""bench.py's CPU leg on the host, without a GPU: the port runner returns the
oracle's state and its own timing, the density norms are the L1 mean and the
maximum, a failing part of the leg is recorded instead of ending the run, and
the host description names the cores this process may use."""
import os
import sys

import numpy as np
import pytest

import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


@pytest.mark.skipif(not os.path.exists(bench.PORT), reason="oracle/_build/fvwb_cpu_fv2d not built")
def test_run_port_equals_oracle_and_times_itself():
    g = 1.3
    p = po.make_params(48, 40, recon=po.RECON_MUSCL_MINMOD, flux=po.FLUX_HLL, bc=po.BC_REFLECT,
                       integrator=po.SSPRK3, gamma=1.4, rho0=1.0, p0=1.0, g=g)
    u, w = po.perturbed_state(p, amp_bump=5e-3, amp_noise=0.0)
    out, sec = bench.run_port(u, w, 6, g, 2)
    uo, _, n, _ = po.fv2d_advance(p, u, w, max_steps=6)
    assert n == 6 and sec >= 0.0
    np.testing.assert_array_equal(out, uo)
    none, sec2 = bench.run_port(u, w, 1, g, 1, want_state=False)
    assert none is None and sec2 >= 0.0


def test_density_err_is_mean_and_max():
    a = np.array([[1.0, 2.0], [3.0, 4.0]])
    b = np.array([[1.5, 2.0], [2.0, 4.0]])
    e = bench.density_err(a, b)
    assert e["l1"] == pytest.approx(0.375) and e["linf"] == 1.0


def test_cpu_leg_records_a_failure():
    def broken():
        raise FileNotFoundError("no such build")
    r = bench.cpu_leg(broken)
    assert "error" in r and "broken" in r["error"] and "no such build" in r["error"]
    assert bench.cpu_leg(lambda x: {"v": x}, 3) == {"v": 3}


def test_host_info_names_the_usable_cores():
    h = bench.host_info()
    assert h["nproc"] == len(os.sched_getaffinity(0)) and h["nproc"] >= 1
    assert h["cpu_count"] == os.cpu_count()
