"""Write the golden fixtures in tests/golden/ from the flang-built reference.

Run in the development container (needs oracle/_ref, built by
``make -C oracle ref`` from the reference's own sources):

    python tests/golden/make_golden.py

Each fixture holds the inputs (or the recipe that regenerates them exactly) and
the reference's outputs.  Large 300x300 outputs are stored as 4096 sampled
entries plus per-variable sums, absolute sums and maxima.  Inputs are rounded
through FP32 so that regenerating them on another CPU (whose exp may differ in
the last FP64 bit) gives the same values.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

import pyoracle as po  # noqa: E402
import ref_runner as rr  # noqa: E402

NSAMPLE = 4096


def bench2d_inputs(salt=1, amp_bump=1e-3, amp_noise=1e-4, nx=300, ny=300):
    """benchmark_2d set-up (nequilibrium=2, 300x300 unless given) plus a
    pressure bump and hash noise; inputs rounded through FP32."""
    p = po.make_params(nx, ny)
    u, w = po.perturbed_state(p, amp_bump=amp_bump, amp_noise=amp_noise, salt=salt)
    u = np.asfortranarray(u.astype(np.float32).astype(np.float64))
    w = np.asfortranarray(w.astype(np.float32).astype(np.float64))
    return p, u, w


def bench1d_inputs(salt=1, nequilibrium=2):
    """benchmark_1d set-up (nequilibrium=2 unless given, nx=128) plus a
    pressure pulse and hash noise on every variable; inputs rounded through FP32."""
    p = po.make_params_1d(128, nequilibrium=nequilibrium)
    u, w = po.perturbed_state_1d(p, amp_bump=1e-3, amp_noise=1e-4, salt=salt)
    u = np.asfortranarray(u.astype(np.float32).astype(np.float64))
    w = np.asfortranarray(w.astype(np.float32).astype(np.float64))
    return p, u, w


def sample_index(shape, n=NSAMPLE, salt=7):
    size = int(np.prod(shape))
    h = po.hash_noise(n, 1, salt)[:, 0]
    return np.unique(((h + 1.0) * 0.5 * size).astype(np.int64) % size)


def summarize(a):
    a = np.asarray(a)
    flat = a.reshape(4, -1, order="F")
    idx = sample_index(a.shape)
    return {
        "idx": idx,
        "sample": a.reshape(-1, order="F")[idx],
        "sum": flat.sum(axis=1),
        "abssum": np.abs(flat).sum(axis=1),
        "maxabs": np.abs(flat).max(axis=1),
    }


def sr_fixture():
    """compute_update_sr (benchmark_1d.f90:553-747, the 'WB1' solver) at the
    reference's nx = 128 on the perturbed state, on its isentropic equilibrium
    (nequilibrium = 2, isothermal, which it does not preserve), on the isentropic
    rest state of ninit = 3 (which the hydrostatic reconstruction preserves) and on
    a state far from both (a jump with flow), so every term is exercised."""
    assert rr.available("bench1d"), "build oracle/_ref first"
    p1, u1, w1 = bench1d_inputs()
    ue = np.asfortranarray(po.conservative_1d(p1.gamma, w1))
    x = (np.arange(p1.nx) + 0.5) / p1.nx
    wj = np.zeros((3, p1.nx), order="F")
    wj[0] = np.where(x < 0.5, 1.0, 0.125) * (1 + 0.1 * np.sin(7 * x))
    wj[1] = 0.3 * np.cos(3 * x)
    wj[2] = np.where(x < 0.5, 1.0, 0.1)
    uj = np.asfortranarray(po.conservative_1d(p1.gamma, wj))
    g = p1.gamma
    wi = np.zeros((3, p1.nx), order="F")
    wi[0] = (1 - ((g - 1) / g) * x) ** (1 / (g - 1))
    wi[2] = (1 - ((g - 1) / g) * x) ** (g / (g - 1))
    ui = np.asfortranarray(po.conservative_1d(g, wi))
    return {"u_isen": ui, "dudt_isen": rr.run("bench1d_update_sr", u=ui, w_eq=w1)["dudt"],
            "u": u1, "w_eq": w1, "dudt": rr.run("bench1d_update_sr", u=u1, w_eq=w1)["dudt"],
            "u_eq": ue, "dudt_eq": rr.run("bench1d_update_sr", u=ue, w_eq=w1)["dudt"],
            "u_jump": uj, "dudt_jump": rr.run("bench1d_update_sr", u=uj, w_eq=w1)["dudt"]}


def fvm1d_fixtures():
    """The two plain 1D residuals: compute_update_fvm (benchmark_1d.f90:454-549,
    solver 'FVM') at nx = 128 on the perturbed and jump states, and compute_update
    of fvm.f90 (:188-251) at its nx = 200 (bc 2, source 2) on its own Sod set-up
    (condinit, ninit = 4), on a smooth moving state and on the polytropic rest
    state of condinit case 7."""
    assert rr.available("bench1d") and rr.available("fvm"), "build oracle/_ref first"
    p1, u1, w1 = bench1d_inputs()
    x = (np.arange(p1.nx) + 0.5) / p1.nx
    wj = np.zeros((3, p1.nx), order="F")
    wj[0] = np.where(x < 0.5, 1.0, 0.125) * (1 + 0.1 * np.sin(7 * x))
    wj[1] = 0.3 * np.cos(3 * x)
    wj[2] = np.where(x < 0.5, 1.0, 0.1)
    uj = np.asfortranarray(po.conservative_1d(p1.gamma, wj))
    a = {"u": u1, "w_eq": w1, "dudt": rr.run("bench1d_update_fvm", u=u1, w_eq=w1)["dudt"],
         "u_jump": uj, "dudt_jump": rr.run("bench1d_update_fvm", u=uj, w_eq=w1)["dudt"]}
    n = 200
    xs = (np.arange(n) + 0.5) / n
    sod = rr.run("fvm_condinit", x=xs)["u"]
    ws = np.zeros((3, n), order="F")
    ws[0] = 1 + 0.2 * np.sin(2 * np.pi * xs)
    ws[1] = 0.5 * np.cos(2 * np.pi * xs)
    ws[2] = 1 + 0.1 * np.cos(4 * np.pi * xs)
    g = p1.gamma
    us = np.asfortranarray(po.conservative_1d(g, ws))
    wp = np.zeros((3, n), order="F")
    wp[0] = 1 - (g - 1) / g * xs
    wp[2] = wp[0] ** g
    up = np.asfortranarray(po.conservative_1d(g, wp))
    b = {"x": xs, "u_sod": sod, "dudt_sod": rr.run("fvm_update", u=sod)["dudt"],
         "u_smooth": us, "dudt_smooth": rr.run("fvm_update", u=us)["dudt"],
         "u_poly": up, "dudt_poly": rr.run("fvm_update", u=up)["dudt"]}
    return {"bench1d_update_fvm_128": a, "fvm_update_200": b}


def dg1d_initial_delta(p, pert=float(np.float32(1e-8))):
    """The reference main program's initial modes for ninit = 8 (dg_with_source.f90:
    168-189): the Gauss-point values of condinit minus the equilibrium,
    projected with its factor 0.5."""
    chsi, wq, ph, _, _ = po.dg1d_tables(p.n)
    dx = p.boxlen / p.nx
    x = ((np.arange(1, p.nx + 1) - 0.5) * dx)[None, :] + dx / 2.0 * chsi[:, None]
    dE = pert * np.exp(-100 * (x - p.boxlen / 2.0) ** 2) / (p.gamma - 1.0)
    d = np.zeros((3, p.n, p.nx), order="F")
    for i in range(p.n):
        d[2, i] = np.sum(0.5 * dE * ph[i][:, None] * wq[:, None], axis=0)
    return d


def dg1d_fixtures():
    """The reference's 1D DG: legendre / legendre_prime (legendre.f90:1-48) for
    n = 0..6 on a grid with the clamp, gl_quadrature for n = 1..6, and
    compute_update_exact_delta (dg_with_source.f90:1749-2031) as compiled (n = 3,
    nx = 128, HLLC, source 2) on its own initial modes, on random modes of size
    1e-2, on strong modes with flow, and on zero modes; and the reference
    program's own run to tend = 0.2 (integrator 'RKi'): every step's t and dt,
    its final error and its recon99999.dat."""
    assert rr.available("dg1d"), "build oracle/_ref first"
    x = np.concatenate([np.linspace(-1, 1, 41), [-1.5, 1.5]])
    r = rr.run("dg1d_legendre", x=x)
    out = {"dg1d_legendre": {"x": x, "P": r["P"], "dP": r["dP"]},
           "dg1d_gl_quadrature": rr.run("dg1d_gl_quadrature")}
    p = po.make_dg1d_params(128)
    ue = po.dg1d_equilibrium_nodes(p)
    rng = np.random.default_rng(1749)
    states = {"init": dg1d_initial_delta(p),
              "rand": np.asfortranarray(rng.normal(size=(3, 3, 128)) * 1e-2),
              "zero": np.zeros((3, 3, 128), order="F")}
    xs = (np.arange(128) + 0.5) / 128
    strong = np.zeros((3, 3, 128), order="F")
    strong[0, 0] = 0.2 * np.sin(6 * xs) * np.exp(-xs)
    strong[1, 0] = 0.3 * np.cos(5 * xs) * np.exp(-xs)
    strong[2, 0] = 0.4 * np.sin(3 * xs) * np.exp(-xs)
    strong[:, 1:] = rng.normal(size=(3, 2, 128)) * 0.02
    states["strong"] = strong
    upd = {"u_eq": ue}
    for k, d in states.items():
        upd["du_" + k] = d
        upd["dudt_" + k] = rr.run("dg1d_update_exact_delta", delta_u=d, u_eq=ue)["dudt"]
    out["dg1d_update_exact_delta_128"] = upd
    assert rr.available("dg1d_program"), "build oracle/_ref first"
    out["dg1d_run_rki_128"] = rr.run_dg1d_program()
    return out


def positivity_fixture():
    """compute_positivity (2d/limiters.f90:478-654, limiter_type 'ONP') as the
    reference compiles it (8x8 cells, mx = my = 2, eps = 1e-10) on tensor modes
    whose cells range from untouched to density- and pressure-limited."""
    assert rr.available("dg2d"), "build oracle/_ref first"
    rng = np.random.default_rng(478)
    nx = ny = 8
    o = 2
    u = np.zeros((4, nx, ny, o, o), order="F")
    rho = rng.uniform(0.5, 1.5, (nx, ny))
    vx, vy = rng.uniform(-1, 1, (nx, ny)), rng.uniform(-1, 1, (nx, ny))
    pr = rng.uniform(0.2, 1.0, (nx, ny))
    u[0, :, :, 0, 0] = rho
    u[1, :, :, 0, 0] = rho * vx
    u[2, :, :, 0, 0] = rho * vy
    u[3, :, :, 0, 0] = pr / 0.4 + 0.5 * rho * (vx ** 2 + vy ** 2)
    scale = rng.choice([0.05, 0.3, 0.8], size=(nx, ny))
    for v in range(4):
        for a in range(o):
            for b in range(o):
                if a or b:
                    u[v, :, :, a, b] = rng.normal(size=(nx, ny)) * scale * np.abs(u[v, :, :, 0, 0] + 0.3)
    r = rr.run("dg2d_positivity", u=u)
    return {"u": u, "u_limited": r["u"], "eps": r["eps"], "gamma": r["gamma"]}


def hllc_fixture():
    """riemann_hllc (dg_with_source.f90:1318-1374) on 1D state pairs chosen to
    reach all four sampled regions (supersonic either way, star states either
    side of the contact) plus equal states and stationary contacts."""
    rng = np.random.default_rng(1318)
    gamma = float(np.float32(1.4))          # the reference's gamma (an FP32 literal)
    m = 384
    ul = np.zeros((m, 3)); ur = np.zeros((m, 3))
    for a in (ul, ur):
        a[:, 0] = rng.uniform(0.2, 2.0, m)
        a[:, 1] = rng.uniform(-3, 3, m) * a[:, 0]
        a[:, 2] = rng.uniform(0.2, 3.0, m) / (gamma - 1) + 0.5 * a[:, 1] ** 2 / a[:, 0]
    ul[:64, 1] = 4.0 * ul[:64, 0]; ur[:64, 1] = 4.0 * ur[:64, 0]            # supersonic to the right
    ul[64:128, 1] = -4.0 * ul[64:128, 0]; ur[64:128, 1] = -4.0 * ur[64:128, 0]  # to the left
    ur[128:160] = ul[128:160]                                              # equal states
    ur[160:192, 1] = 0.0; ul[160:192, 1] = 0.0                             # stationary contacts
    ur[160:192, 2] = ul[160:192, 2]
    for a in (ul, ur):
        a[:, 2] = np.maximum(a[:, 2], 0.2 / (gamma - 1) + 0.5 * a[:, 1] ** 2 / a[:, 0])
    f = rr.run("dg1d_riemann_hllc", ul=ul, ur=ur, gamma=np.array(gamma))["f"]
    return {"ul": ul, "ur": ur, "f": f, "gamma": np.array(gamma)}


def plain_fixture():
    """compute_update (benchmark_2d.f90:370-463), the plain non-WB residual, on the
    perturbed 300x300 input of the compute_update_exact fixture."""
    p, u, w = bench2d_inputs(salt=3, amp_bump=1e-2, amp_noise=1e-3)
    d = rr.run("bench2d_update_plain", u=u, w_eq=w)["dudt"]
    return dict(summarize(d), salt=3, amp_bump=1e-2, amp_noise=1e-3)


# the reference's 2D DG builds: the shipped set-up and three edits of its
# compile-time switches (oracle/Makefile); the oracle's parameters for each
DGQ_BUILDS = {
    "dg2d": {},
    "dgq_hll2": dict(order=3, flux_type=po.DGQ_FLUX_HLL2, source=2, solver=po.DGQ_SOLVER_DEB),
    "dgq_llf1": dict(flux_type=po.DGQ_FLUX_LLF1, source=2, grad_phi_case=1, bc=2, ninit=2,
                     solver=po.DGQ_SOLVER_SS4),
    "dgq_hllc": dict(order=3, flux_type=po.DGQ_FLUX_HLLC, source=3, bc=3, ninit=12, solver=po.DGQ_SOLVER_EQL),
}
DGQ_TEND = {"dg2d": 0.05, "dgq_hll2": 0.02, "dgq_llf1": 0.05, "dgq_hllc": 0.05}


def dgq_perturbed_modes(lib, u0, o, seed=5):
    """modes of the build's own initial nodal state with noise on every
    variable (projected by the reference's get_modes_from_nodes)"""
    rng = np.random.default_rng(seed)
    nodes = u0.copy(order="F")
    nodes[0] *= 1 + 0.05 * rng.standard_normal(nodes[0].shape)
    nodes[1] += 0.1 * rng.standard_normal(nodes[0].shape)
    nodes[2] += 0.1 * rng.standard_normal(nodes[0].shape)
    nodes[3] *= 1 + 0.05 * rng.random(nodes[0].shape)
    nodes = np.asfortranarray(nodes.astype(np.float32).astype(np.float64))
    return rr.run("dg2d_modes_from_nodes", a=nodes)["r"] if lib == "dg2d" else _modes(nodes, o)


def _modes(nodes, o):
    m = np.zeros_like(nodes, order="F")
    po.lib().orc_ref_dg2d_modes_from_nodes(nodes.shape[1], nodes.shape[2], o, po._p(nodes), po._p(m))
    return m


def dgq_fixtures():
    """The reference's own 2D DG program, per build: its set-up (get_coords,
    get_initial_conditions, get_equilibrium_solution), compute_update on a
    perturbed state, compute_max_speed of its means, apply_limiter, and a whole
    evolve run from its initial condition (every dt, the final nodal state,
    compute_error's printed norms, and the snapshot files output_file wrote)."""
    out = {}
    for lib, kw in DGQ_BUILDS.items():
        p = po.make_dgq_params(8, **kw)
        o = p.order
        r = rr.run("dgq_setup", lib=np.array(lib), n=np.array(8), o=np.array(o))
        x, y, u0 = r["x"], r["y"], r["u"]
        du = dgq_perturbed_modes(lib, u0, o)
        d = rr.run("dgq_update", lib=np.array(lib), delta_u=du, x=x, y=y)["dudt"]
        f = {"x": x, "y": y, "u0": u0, "u_eq": r["u_eq"], "delta_u": du, "dudt": d}
        if p.ninit == 12:       # coordinates scaled so that special_boundary_conditions freezes cells
            f["dudt_x6"] = rr.run("dgq_update", lib=np.array(lib), delta_u=du, x=6 * x, y=6 * y)["dudt"]
        means = np.asfortranarray(du[:, :, :, 0, 0])
        f["max_speed"] = rr.run("dgq_max_speed", lib=np.array(lib), u=means)["out"]
        lim_in = du.copy(order="F")
        lim_in[:, :, :, 1:, :] *= 40.0        # steep enough that density and pressure are limited
        lim_in[:, :, :, 0, 1:] *= 40.0
        f["limiter_in"] = lim_in
        f["limiter_out"] = rr.run("dgq_limiter", lib=np.array(lib), u=lim_in)["u"]
        ev = rr.run_dg2d_evolve(lib, u0, x, y, DGQ_TEND[lib])
        f.update(tend=np.array(DGQ_TEND[lib]), evolve_u=ev["u"], evolve_t=ev["t"], evolve_dt=ev["dt"],
                 evolve_iters=ev["iters"], norms=ev["norms"], norm_keys=ev["norm_keys"], files=ev["files"])
        for k in ev:
            if k.startswith("file_"):
                f[k] = ev[k]
        out["dgq_" + ("shipped" if lib == "dg2d" else lib[4:])] = f
    return out


def fv_variant_fixtures():
    """The edited FV set-ups: compute_update_exact, compute_max_speed and three
    Heun steps of evolve at 300 x 200 (dx != dy), and the 1D compute_update and
    compute_max_speed with the polytropic equilibrium (nequilibrium 3)."""
    assert rr.available("bench2d_300x200") and rr.available("bench1d_neq3"), "build oracle/_ref first"
    out = {}
    p, u, w = bench2d_inputs(salt=2, nx=300, ny=200)
    lib = np.array("bench2d_300x200")
    d = rr.run("bench2d_update_exact", u=u, w_eq=w, lib=lib)["dudt"]
    c = rr.run("bench2d_max_speed", u=u, lib=lib)["cmax"]
    tend = 2.5 * (0.5 * (1.0 / 300) / float(c[0]) * 0.5)
    ueq = po.conservative(p.gamma, w)
    ue = rr.run("bench2d_evolve", u=u, w_eq=w, tend=np.array(tend), lib=lib)["u"]
    out["bench2d_update_exact_300x200"] = dict(summarize(d), salt=2, cmax=c)
    out["bench2d_evolve_300x200"] = dict(summarize(ue - ueq), tend=np.array(tend), salt=2)
    p1, u1, w1 = bench1d_inputs(salt=3, nequilibrium=3)
    lib1 = np.array("bench1d_neq3")
    ue1 = po.conservative_1d(p1.gamma, w1)
    out["bench1d_update_neq3_128"] = {
        "u": u1, "w_eq": w1, "dudt": rr.run("bench1d_update", u=u1, w_eq=w1, lib=lib1)["dudt"],
        "dudt_eq": rr.run("bench1d_update", u=ue1, w_eq=w1, lib=lib1)["dudt"],
        "cmax": rr.run("bench1d_max_speed", u=u1, lib=lib1)["cmax"]}
    return out


def main(only=None):
    assert rr.available("bench2d") and rr.available("dg2d"), "build oracle/_ref first"
    out = {}
    single = {"dg1d_riemann_hllc": hllc_fixture, "bench2d_update_plain_300": plain_fixture,
              "bench1d_update_sr_128": sr_fixture, "dg2d_positivity": positivity_fixture}
    if only in single:
        out[only] = single[only]()
    elif only == "fvm1d":
        out.update(fvm1d_fixtures())
    elif only == "dg1d":
        out.update(dg1d_fixtures())
    elif only == "dgq":
        out.update(dgq_fixtures())
    elif only == "fvvar":
        out.update(fv_variant_fixtures())
    if out:
        for name, arrs in out.items():
            path = os.path.join(HERE, name + ".npz")
            np.savez_compressed(path, **{k: np.asarray(v) for k, v in arrs.items()})
            print(f"wrote {path} ({os.path.getsize(path)} bytes)")
        return

    # A1 compute_update_exact on a perturbed state and at equilibrium
    p, u, w = bench2d_inputs()
    d = rr.run("bench2d_update_exact", u=u, w_eq=w)["dudt"]
    out["bench2d_update_exact_300"] = dict(summarize(d), salt=1, amp_bump=1e-3, amp_noise=1e-4)
    ueq = po.conservative(p.gamma, w)
    d0 = rr.run("bench2d_update_exact", u=ueq, w_eq=w)["dudt"]
    out["bench2d_update_exact_eq_300"] = {"maxabs": np.abs(d0).max(axis=(1, 2)),
                                          "nonzero": np.array(np.count_nonzero(d0))}

    # A7 compute_max_speed
    c = rr.run("bench2d_max_speed", u=u)["cmax"]
    out["bench2d_max_speed_300"] = {"cmax": c, "salt": 1}

    # A8 evolve: tend = 2.5 * first dt -> three Heun steps
    dt0 = 0.5 * (1.0 / 300) / float(c[0]) * 0.5
    tend = 2.5 * dt0
    ue = rr.run("bench2d_evolve", u=u, w_eq=w, tend=np.array(tend))["u"]
    out["bench2d_evolve_300"] = dict(summarize(ue - ueq), tend=np.array(tend), salt=1)

    # A10 minmod / TVB generalized_minmod (2d/limiters.f90:18-54; dx = 1/8)
    rng = np.random.default_rng(20261202)
    xyz = rng.normal(size=(512, 3))
    xyz[::7, 1] = 0.0
    xyz[::11, 0] = 0.0
    xyz[::13, 2] = -0.0
    r = rr.run("dg2d_minmod", xyz=xyz, M=np.array(50.0))
    out["minmod"] = {"xyz": xyz, "M": np.array(50.0), "dx": np.array(1.0 / 8),
                     "minmod": r["minmod"], "generalized_minmod": r["generalized_minmod"]}

    # A11 compute_hllflux (as written) and states for flux consistency
    n = 256
    ul = np.zeros((n, 4)); ur = np.zeros((n, 4))
    for a in (ul, ur):
        a[:, 0] = rng.uniform(0.2, 2.0, n)
        a[:, 1] = rng.uniform(-2, 2, n) * a[:, 0]
        a[:, 2] = rng.uniform(-2, 2, n) * a[:, 0]
        a[:, 3] = rng.uniform(0.2, 3.0, n) / (p.gamma - 1) + 0.5 * (a[:, 1] ** 2 + a[:, 2] ** 2) / a[:, 0]
    flag = rng.integers(1, 3, n).astype(np.int32)
    fl = np.array([po.ref_flux(p.gamma, x)[f - 1] for x, f in zip(ul, flag)])
    fr = np.array([po.ref_flux(p.gamma, x)[f - 1] for x, f in zip(ur, flag)])
    f = rr.run("dg2d_hllflux", ul=ul, ur=ur, fl=fl, fr=fr, flag=flag)["f"]
    out["dg2d_hllflux"] = {"ul": ul, "ur": ur, "fl": fl, "fr": fr, "flag": flag, "f": f,
                           "gamma": np.array(p.gamma)}

    # A18 Legendre basis and GL quadrature (2d/legendre.f90)
    x = np.concatenate([np.linspace(-1, 1, 41), [-1.5, 1.5]])
    r = rr.run("dg2d_legendre", x=x)
    out["dg2d_legendre"] = {"x": x, "P": r["P"], "dP": r["dP"]}
    out["dg2d_gl_quadrature"] = rr.run("dg2d_gl_quadrature")

    # A13 benchmark_1d compute_update ('EQL') and compute_max_speed, nx = 128
    assert rr.available("bench1d"), "build oracle/_ref first"
    p1, u1, w1 = bench1d_inputs()
    d1 = rr.run("bench1d_update", u=u1, w_eq=w1)["dudt"]
    c1 = rr.run("bench1d_max_speed", u=u1)["cmax"]
    ue1 = po.conservative_1d(p1.gamma, w1)
    d0 = rr.run("bench1d_update", u=ue1, w_eq=w1)["dudt"]
    out["bench1d_update_128"] = {"u": u1, "w_eq": w1, "dudt": d1, "cmax": c1, "dudt_eq": d0}
    # compute_llflux on random state pairs
    m = 256
    ul = np.zeros((m, 3)); ur = np.zeros((m, 3))
    for a in (ul, ur):
        a[:, 0] = rng.uniform(0.2, 2.0, m)
        a[:, 1] = rng.uniform(-2, 2, m) * a[:, 0]
        a[:, 2] = rng.uniform(0.2, 3.0, m) / (p1.gamma - 1) + 0.5 * a[:, 1] ** 2 / a[:, 0]
    fl1 = np.array([po.ref1d_flux(p1.gamma, x) for x in ul])
    fr1 = np.array([po.ref1d_flux(p1.gamma, x) for x in ur])
    f1 = rr.run("bench1d_llflux", ul=ul, ur=ur, fl=fl1, fr=fr1)["f"]
    out["bench1d_llflux"] = {"ul": ul, "ur": ur, "fl": fl1, "fr": fr1, "f": f1, "gamma": np.array(p1.gamma)}

    # compute_update_sr, the 'WB1' hydrostatic reconstruction
    out["bench1d_update_sr_128"] = sr_fixture()

    # the plain 1D residuals: benchmark_1d 'FVM' and fvm.f90
    out.update(fvm1d_fixtures())

    # the reference's shipped 2D DG limiter, compute_positivity
    out["dg2d_positivity"] = positivity_fixture()

    # the reference's 1D DG: basis, quadrature, compute_update_exact_delta
    out.update(dg1d_fixtures())

    # A9 compute_update, the plain non-WB residual
    out["bench2d_update_plain_300"] = plain_fixture()

    # dg_with_source.f90 riemann_hllc (the HLLC of the DG codes)
    out["dg1d_riemann_hllc"] = hllc_fixture()

    # A17 tensor-basis projections of the reference's 2D DG (8x8 cells, order 2)
    a = np.asfortranarray(np.random.default_rng(17).normal(size=(4, 8, 8, 2, 2)))
    out["dg2d_projections"] = {"a": a, "modes_from_nodes": rr.run("dg2d_modes_from_nodes", a=a)["r"],
                               "nodes_from_modes": rr.run("dg2d_nodes_from_modes", a=a)["r"]}

    # the reference's own 2D DG program, shipped and with edited switches
    out.update(dgq_fixtures())

    # the FV set-ups at nx != ny and with nequilibrium 3
    out.update(fv_variant_fixtures())

    for name, arrs in out.items():
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **{k: np.asarray(v) for k, v in arrs.items()})
        print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else None)
