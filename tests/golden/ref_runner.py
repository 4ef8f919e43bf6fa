"""Run subroutines of the flang-built reference (oracle/_ref/libref_*.so).

Used only as a checker: by make_golden.py to write the committed fixtures and by
the CPU tests when oracle/_ref is present.  Every call runs in a child process
(``run(op, **arrays)``) because the reference prints whole arrays to stdout on
every call (benchmark_2d.f90:610) and keeps ~30 full-size automatic arrays on the
stack (benchmark_2d.f90:468-483): the child sends stdout to /dev/null and runs the
call on a thread with a 2 GiB stack.

The reference's grid sizes are compile-time parameters of its own modules:
benchmark_2d is 300x300 (parameters_2d.f90:3-4), benchmark_1d has nx=128
(parameters.f90:3), fvm.f90 has nx=200, bc=2, source=2 and the Sod set-up
ninit=4 (fvm_commons.f90:5-14), the 1D DG has n=3 modes on nx=128 cells with
HLLC and the gravity source (dg_commons.f90:4-15), 2d DG is 8x8 cells with mx=my=2 (2d/parameters_dg_2d.f90:3-6).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import sys
import tempfile
import threading

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

LIBS = {
    "bench2d": "libref_bench2d.so",
    "bench1d": "libref_bench1d.so",
    # edits of the FV set-ups (oracle/Makefile): nx != ny, nequilibrium 3
    "bench2d_300x200": "libref_bench2d_300x200.so",
    "bench1d_neq3": "libref_bench1d_neq3.so",
    "dg2d": "libref_dg2d.so",
    "dg1d": "libref_dg1d.so",
    "fvm": "libref_fvm.so",
    # variants of the 2D DG set-up's compile-time switches (oracle/Makefile)
    "dgq_hll2": "libref_dgq_hll2.so",
    "dgq_llf1": "libref_dgq_llf1.so",
    "dgq_hllc": "libref_dgq_hllc.so",
}


def available(name="bench2d") -> bool:
    if name == "dg1d_program":
        return os.path.exists(os.path.join(REF_DIR, "ref_dg1d"))
    return os.path.exists(os.path.join(REF_DIR, LIBS[name]))


def run(op: str, **arrays):
    """Run `op` in a child process; returns a dict of numpy arrays."""
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.npz"), os.path.join(td, "out.npz")
        np.savez(fin, **{k: np.asarray(v) for k, v in arrays.items()})
        with open(os.devnull, "w") as dn:
            subprocess.run([sys.executable, os.path.abspath(__file__), op, fin, fout],
                           stdout=dn, check=True)
        with np.load(fout) as z:
            return {k: z[k] for k in z.files}


def run_dg2d_evolve(lib: str, u, x, y, tend):
    """Call evolve (2d/benchmark_2d_dg.f90:624-775) of the 2D DG build `lib`
    ("dg2d" for the shipped set-up, or a variant) on the nodal state u with
    tend set in its module, in a child process whose working directory holds
    the longrotation2_o1/ folder that output_file writes into and whose stdin
    is not a terminal (so the two `pause` statements do not wait).  Returns
    the final nodal state, the (iter, t, dt) of every step from its
    'time=' lines (:746), compute_error's printed norms (:38-88), and the
    snapshot files it wrote with their names."""
    import re
    with tempfile.TemporaryDirectory() as td:
        os.makedirs(os.path.join(td, "longrotation2_o1"))
        fin, fout = os.path.join(td, "in.npz"), os.path.join(td, "out.npz")
        np.savez(fin, u=np.asarray(u), x=np.asarray(x), y=np.asarray(y), tend=np.array(float(tend)),
                 lib=np.array(lib))
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "dgq_evolve", fin, fout], cwd=td,
                           stdin=subprocess.DEVNULL, capture_output=True, text=True, check=True, timeout=1800)
        with np.load(fout) as z:
            out = {k: z[k] for k in z.files}
        rows = re.findall(r"time=\s+(\d+)\s+(\S+)\s+(\S+)", r.stdout)
        out["iters"] = np.array([int(a) for a, _, _ in rows])
        out["t"] = np.array([float(b) for _, b, _ in rows])
        out["dt"] = np.array([float(c) for _, _, c in rows])
        norms = {}
        for key in ("maxerror rho", "maxerror velx", "maxerror vely", "maxerror energy", "l1error rho",
                    "l1error velx", "l1error vely", "l1error energy", "l2error rho", "l2error velx",
                    "l2error vely", "l2error energy"):
            m = re.findall(re.escape(key) + r"\s+(\S+)", r.stdout)
            norms[key] = float(m[-1])
        out["norms"] = np.array([norms[k] for k in sorted(norms)])
        out["norm_keys"] = np.array(sorted(norms))
        files = sorted(os.listdir(os.path.join(td, "longrotation2_o1")))
        out["files"] = np.array(files)
        for f in files:
            with open(os.path.join(td, "longrotation2_o1", f)) as fh:
                out["file_" + f] = np.array(fh.read())
    return out


def run_dg1d_program():
    """Run the reference's 1D DG program (oracle/_ref/ref_dg1d, dg_commons.f90's
    set-up: n = 3, nx = 128, HLLC, ninit = 8, integrator 'RKi', tend = 0.2) in an
    empty directory and return what it reports: the time and dt of every step
    (its stdout, :196-199 and :333), the final 'error' (:406), and the two columns
    of recon99999.dat (:341-349: the first Gauss point's x and p - p_eq there)."""
    import re
    with tempfile.TemporaryDirectory() as td:
        os.makedirs(os.path.join(td, "simul"))
        r = subprocess.run([os.path.join(REF_DIR, "ref_dg1d")], cwd=td, capture_output=True, text=True,
                           check=True, timeout=600)
        rows = re.findall(r"time=\s+(\d+)\s+(\S+)\s+(\S+)", r.stdout)
        err = re.findall(r"error\s+(\S+)", r.stdout)
        recon = np.loadtxt(os.path.join(td, "recon99999.dat"))
    return {"steps": np.array([int(a) for a, _, _ in rows]), "t": np.array([float(b) for _, b, _ in rows]),
            "dt": np.array([float(c) for _, _, c in rows]), "error": np.array(float(err[-1])), "recon": recon}


# ---- child side ------------------------------------------------------------------
def _dp(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _child(op, fin, fout):
    z = dict(np.load(fin))
    out = {}
    if op.startswith("bench2d_"):
        L = ctypes.CDLL(os.path.join(REF_DIR, LIBS[str(z["lib"]) if "lib" in z else "bench2d"]))
        if op == "bench2d_update_exact":
            u = np.asfortranarray(z["u"]); w = np.asfortranarray(z["w_eq"])
            d = np.zeros_like(u, order="F")
            L.compute_update_exact_(_dp(u), _dp(w), _dp(d))
            out["dudt"] = d
        elif op == "bench2d_update_plain":
            u = np.asfortranarray(z["u"]); w = np.asfortranarray(z["w_eq"])
            d = np.zeros_like(u, order="F")
            L.compute_update_(_dp(u), _dp(w), _dp(d))
            out["dudt"] = d
        elif op == "bench2d_max_speed":
            u = np.asfortranarray(z["u"])
            c = np.zeros(1)
            L.compute_max_speed_(_dp(u), _dp(c))
            out["cmax"] = c
        elif op == "bench2d_evolve":
            u = np.asfortranarray(z["u"]).copy(order="F"); w = np.asfortranarray(z["w_eq"])
            tend = ctypes.c_double.in_dll(L, "_QMparameters_2dEtend")
            tend.value = float(z["tend"])
            L.evolve_(_dp(u), _dp(w))
            out["u"] = u
        elif op == "bench2d_equilibrium":
            x = np.asfortranarray(z["x"]); y = np.asfortranarray(z["y"])
            n1, n2 = x.shape
            w = np.zeros((4, n1, n2), order="F")
            a, b = ctypes.c_int(n1), ctypes.c_int(n2)
            L.get_equilibrium_solution_(_dp(x), _dp(y), _dp(w), ctypes.byref(a), ctypes.byref(b))
            out["w"] = w
        else:
            raise ValueError(op)
    elif op.startswith("bench1d_"):
        L = ctypes.CDLL(os.path.join(REF_DIR, LIBS[str(z["lib"]) if "lib" in z else "bench1d"]))
        if op == "bench1d_update":
            u = np.asfortranarray(z["u"]); w = np.asfortranarray(z["w_eq"])
            d = np.zeros_like(u, order="F")
            L.compute_update_(_dp(u), _dp(w), _dp(d))
            out["dudt"] = d
        elif op == "bench1d_update_sr":
            u = np.asfortranarray(z["u"]); w = np.asfortranarray(z["w_eq"])
            d = np.zeros_like(u, order="F")
            L.compute_update_sr_(_dp(u), _dp(w), _dp(d))
            out["dudt"] = d
        elif op == "bench1d_update_fvm":
            u = np.asfortranarray(z["u"]); w = np.asfortranarray(z["w_eq"])
            d = np.zeros_like(u, order="F")
            L.compute_update_fvm_(_dp(u), _dp(w), _dp(d))
            out["dudt"] = d
        elif op == "bench1d_max_speed":
            u = np.asfortranarray(z["u"])
            c = np.zeros(1)
            L.compute_max_speed_(_dp(u), _dp(c))
            out["cmax"] = c
        elif op == "bench1d_llflux":
            ul, ur, fl, fr = (np.ascontiguousarray(z[k]) for k in ("ul", "ur", "fl", "fr"))
            res = np.zeros_like(ul)
            for q in range(ul.shape[0]):
                o = np.zeros(3)
                a, b, c, e = (np.ascontiguousarray(v[q]) for v in (ul, ur, fl, fr))
                L.compute_llflux_(_dp(a), _dp(b), _dp(c), _dp(e), _dp(o))
                res[q] = o
            out["f"] = res
        else:
            raise ValueError(op)
    elif op.startswith("fvm_"):
        L = ctypes.CDLL(os.path.join(REF_DIR, LIBS["fvm"]))
        if op == "fvm_update":
            u = np.asfortranarray(z["u"])
            d = np.zeros_like(u, order="F")
            L.compute_update_(_dp(u), _dp(d))
            out["dudt"] = d
        elif op == "fvm_condinit":
            x = np.ascontiguousarray(z["x"])
            res = np.zeros((3, len(x)), order="F")
            for i, xv in enumerate(x):
                xx, uu = ctypes.c_double(float(xv)), np.zeros(3)
                L.condinit_(ctypes.byref(xx), _dp(uu))
                res[:, i] = uu
            out["u"] = res
        else:
            raise ValueError(op)
    elif op.startswith("dg1d_"):
        L = ctypes.CDLL(os.path.join(REF_DIR, LIBS["dg1d"]))
        if op == "dg1d_legendre":
            x = np.ascontiguousarray(z["x"])
            L.legendre_.restype = ctypes.c_double
            L.legendre_prime_.restype = ctypes.c_double
            P = np.zeros((7, len(x))); dP = np.zeros((7, len(x)))
            for k in range(7):
                for i, xv in enumerate(x):
                    a, b = ctypes.c_double(float(xv)), ctypes.c_double(float(xv))
                    P[k, i] = L.legendre_(ctypes.byref(a), ctypes.byref(ctypes.c_int(k)))
                    dP[k, i] = L.legendre_prime_(ctypes.byref(b), ctypes.byref(ctypes.c_int(k)))
            out["P"], out["dP"] = P, dP
        elif op == "dg1d_gl_quadrature":
            for n in range(1, 7):
                xq, wq = np.zeros(n), np.zeros(n)
                L.gl_quadrature_(_dp(xq), _dp(wq), ctypes.byref(ctypes.c_int(n)))
                out[f"x{n}"], out[f"w{n}"] = xq, wq
        elif op == "dg1d_update_exact_delta":
            d = np.asfortranarray(z["delta_u"]); e = np.asfortranarray(z["u_eq"])
            r = np.zeros_like(d, order="F")
            L.compute_update_exact_delta_(_dp(d), _dp(e), _dp(r))
            out["dudt"] = r
        elif op == "dg1d_riemann_hllc":
            ul, ur = (np.ascontiguousarray(z[k]) for k in ("ul", "ur"))
            gam = ctypes.c_double(float(z["gamma"]))
            nv = ctypes.c_int(3)
            res = np.zeros_like(ul)
            for q in range(ul.shape[0]):
                o = np.zeros(3)
                a, b = np.ascontiguousarray(ul[q]), np.ascontiguousarray(ur[q])
                L.riemann_hllc_(_dp(a), _dp(b), _dp(o), ctypes.byref(gam), ctypes.byref(nv))
                res[q] = o
            out["f"] = res
        else:
            raise ValueError(op)
    elif op.startswith("dgq_"):
        L = ctypes.CDLL(os.path.join(REF_DIR, LIBS[str(z["lib"])]))
        mod = "_QMparameters_dg_2dE"
        if op == "dgq_update":
            du = np.asfortranarray(z["delta_u"]); x = np.asfortranarray(z["x"]); y = np.asfortranarray(z["y"])
            ueq = np.zeros_like(du, order="F")
            d = np.zeros_like(du, order="F")
            L.compute_update_(_dp(du), _dp(x), _dp(y), _dp(ueq), _dp(d))
            out["dudt"] = d
        elif op == "dgq_max_speed":
            u = np.asfortranarray(z["u"])
            r = [ctypes.c_double() for _ in range(4)]
            L.compute_max_speed_(_dp(u), *[ctypes.byref(v) for v in r])
            out["out"] = np.array([v.value for v in r])
        elif op == "dgq_limiter":
            u = np.asfortranarray(z["u"]).copy(order="F")
            L.apply_limiter_(_dp(u))
            out["u"] = u
        elif op == "dgq_setup":
            n, o = int(z["n"]), int(z["o"])
            x = np.zeros((n, n, o, o), order="F"); y = np.zeros((n, n, o, o), order="F")
            u = np.zeros((4, n, n, o, o), order="F"); ueq = np.zeros((4, n, n, o, o), order="F")
            sx, sy, ox, oy = (ctypes.c_int(v) for v in (n, n, o, o))
            args = (ctypes.byref(sx), ctypes.byref(sy), ctypes.byref(ox), ctypes.byref(oy))
            L.get_coords_(_dp(x), _dp(y), *args)
            L.get_initial_conditions_(_dp(x), _dp(y), _dp(u), *args)
            L.get_equilibrium_solution_(_dp(x), _dp(y), _dp(ueq), *args)
            out.update(x=x, y=y, u=u, u_eq=ueq)
        elif op == "dgq_evolve":
            u = np.asfortranarray(z["u"]).copy(order="F")
            x = np.asfortranarray(z["x"]); y = np.asfortranarray(z["y"])
            ueq = np.zeros_like(u, order="F")
            ctypes.c_double.in_dll(L, mod + "tend").value = float(z["tend"])
            L.evolve_(_dp(u), _dp(x), _dp(y), _dp(ueq))
            sys.stdout.flush()
            out["u"] = u
        else:
            raise ValueError(op)
    elif op.startswith("dg2d_"):
        L = ctypes.CDLL(os.path.join(REF_DIR, LIBS["dg2d"]))
        d = ctypes.c_double
        if op == "dg2d_minmod":
            L.minmod_.restype = d
            L.generalized_minmod_.restype = d
            xs = z["xyz"]
            mm = np.zeros(len(xs)); gm = np.zeros(len(xs))
            Mv = d.in_dll(L, "_QMparameters_dg_2dEm")
            Mv.value = float(z["M"])
            for q, (a, b, c) in enumerate(xs):
                aa, bb, cc = d(a), d(b), d(c)
                mm[q] = L.minmod_(ctypes.byref(aa), ctypes.byref(bb), ctypes.byref(cc))
                aa, bb, cc = d(a), d(b), d(c)
                gm[q] = L.generalized_minmod_(ctypes.byref(aa), ctypes.byref(bb), ctypes.byref(cc))
            out["minmod"] = mm
            out["generalized_minmod"] = gm
        elif op == "dg2d_hllflux":
            ul, ur, fl, fr = (np.ascontiguousarray(z[k]) for k in ("ul", "ur", "fl", "fr"))
            n = ul.shape[0]
            res = np.zeros((n, 4))
            for q in range(n):
                o = np.zeros(4)
                flag = ctypes.c_int(int(z["flag"][q]))
                a, b, c, e = (np.ascontiguousarray(v[q]) for v in (ul, ur, fl, fr))
                L.compute_hllflux_(_dp(a), _dp(b), _dp(c), _dp(e), _dp(o), ctypes.byref(flag))
                res[q] = o
            out["f"] = res
        elif op == "dg2d_legendre":
            L.legendre_.restype = d
            L.legendre_prime_.restype = d
            xs = z["x"]
            P = np.zeros((len(xs), 7)); dP = np.zeros((len(xs), 7))
            for q, xv in enumerate(xs):
                for n in range(7):
                    a = d(xv); nn = ctypes.c_int(n)
                    P[q, n] = L.legendre_(ctypes.byref(a), ctypes.byref(nn))
                    a = d(xv)
                    dP[q, n] = L.legendre_prime_(ctypes.byref(a), ctypes.byref(nn))
            out["P"] = P
            out["dP"] = dP
        elif op in ("dg2d_modes_from_nodes", "dg2d_nodes_from_modes"):
            a = np.asfortranarray(z["a"])
            _, nx, ny, o, _ = a.shape
            res = np.zeros_like(a, order="F")
            fn = L.get_modes_from_nodes_ if op == "dg2d_modes_from_nodes" else L.get_nodes_from_modes_
            sx, sy, ox, oy = (ctypes.c_int(v) for v in (nx, ny, o, o))
            fn(_dp(a), _dp(res), ctypes.byref(sx), ctypes.byref(sy), ctypes.byref(ox), ctypes.byref(oy))
            out["r"] = res
        elif op == "dg2d_positivity":
            u = np.asfortranarray(z["u"]).copy(order="F")
            L.compute_positivity_(_dp(u))
            out["u"] = u
            out["eps"] = np.array(ctypes.c_double.in_dll(L, "_QMparameters_dg_2dEeps").value)
            out["gamma"] = np.array(ctypes.c_double.in_dll(L, "_QMparameters_dg_2dEgamma").value)
        elif op == "dg2d_gl_quadrature":
            res = {}
            for n in range(1, 7):
                xq = np.zeros(n); wq = np.zeros(n)
                nn = ctypes.c_int(n)
                L.gl_quadrature_(_dp(xq), _dp(wq), ctypes.byref(nn))
                out[f"x{n}"] = xq
                out[f"w{n}"] = wq
        else:
            raise ValueError(op)
    else:
        raise ValueError(op)
    np.savez(fout, **out)


if __name__ == "__main__":
    threading.stack_size(1 << 31)
    err = []

    def _target():
        try:
            _child(*sys.argv[1:4])
        except BaseException as e:  # noqa: BLE001 - reported to the parent below
            err.append(e)

    th = threading.Thread(target=_target)
    th.start()
    th.join()
    if err:
        sys.stderr.write(f"ref_runner {sys.argv[1]} failed: {err[0]!r}\n")
        sys.exit(1)
