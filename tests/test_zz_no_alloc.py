"""fvwb_dgq_compute_error_noalloc, the allocation-free variant of
fvwb_dgq_compute_error (compute_error, 2d/benchmark_2d_dg.f90:23-89): on
device arrays the caller owns, with its block partial sums in the library's
static scratch array, it returns the handle entry point's norms bit for bit;
two calls in a row, the second on a smaller grid, so stale partials of the
first call would show; and sizes beyond the scratch array are refused."""
import ctypes
import sys

import numpy as np
import pytest

import fvwb

pytestmark = pytest.mark.gpu


class Hip:
    """The caller's own device memory, through the HIP runtime."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")
        for n in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize"):
            getattr(self.h, n).restype = ctypes.c_int
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]

    def put(self, a):
        a = np.asfortranarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        return p

    def zeros(self, n):
        return self.put(np.zeros(n))

    def get(self, p, n):
        out = np.zeros(n)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0
        return out

    def free(self, *ps):
        for p in ps:
            assert self.h.hipFree(p) == 0


def test_noalloc_error_equals_handle_entry_point():
    hip = Hip()
    L = fvwb.lib()
    for n, order, seed in ((96, 3, 1), (24, 2, 2)):
        print(f"case: compute_error_noalloc {n}x{n} order {order}", flush=True)
        sys.stdout.flush()
        rng = np.random.default_rng(seed)
        shape = (4, n, n, order, order)
        nodes = np.asfortranarray(rng.standard_normal(shape))
        u_init = np.asfortranarray(nodes + 1e-3 * rng.standard_normal(shape))
        with fvwb.SolverDGQ(fvwb.reference_params_dgq(n, order=order)) as s:
            linf, l1, l2 = s.compute_error(u_init, nodes)
        dn, du, do = hip.put(nodes), hip.put(u_init), hip.zeros(12)
        rc = L.fvwb_dgq_compute_error_noalloc(n, n, order, dn, du, do, None)
        assert rc == fvwb.OK, L.fvwb_dgq_last_error(None)
        got = hip.get(do, 12)
        hip.free(dn, du, do)
        np.testing.assert_array_equal(got[:4], linf)
        np.testing.assert_array_equal(got[4:8], l1)
        np.testing.assert_array_equal(got[8:], l2)
        assert np.all(got[4:] > 0)


def test_noalloc_error_refuses_sizes_beyond_its_scratch():
    L = fvwb.lib()
    one = ctypes.c_void_p(16)     # never dereferenced: the size check comes first
    rc = L.fvwb_dgq_compute_error_noalloc(1025, 1024, 2, one, one, one, None)
    assert rc == fvwb.ERR_ARG
    assert b"FVWB_DGQ_NOALLOC_MAX_CELLS" in L.fvwb_dgq_last_error(None)
    assert L.fvwb_dgq_compute_error_noalloc(8, 8, 4, one, one, one, None) == fvwb.ERR_ARG
    assert L.fvwb_dgq_compute_error_noalloc(8, 8, 2, None, one, one, None) == fvwb.ERR_ARG
