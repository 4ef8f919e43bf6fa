/*
 * fvwb.h — C ABI of the MI355X well-balanced Euler-with-gravity solver.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, returns an
 * int status (FVWB_OK == 0, negative on error) and never throws.  Host arrays use
 * the reference's Fortran layout: a conservative state `u(1:nvar,1:nx,1:ny)` is a
 * contiguous double array with the variable index fastest, then x, then y
 * (element (v,i,j) at offset v + nvar*(i + nx*j), zero-based).
 *
 * Physics: 2D Euler, gamma-law gas, linear potential phi = g*(x+y) on
 * [0,boxlen_x]x[0,boxlen_y].  The stored hydrostatic equilibrium is the
 * isothermal atmosphere rho = rho0*exp(-k*(x+y)), p = p0*exp(-k*(x+y)),
 * k = rho0*g/p0, at rest (benchmark_2d.f90:174-218, get_equilibrium_solution).
 * The device state is the deviation dU = U - U_eq, one row per variable and grid
 * row (the four rows of a grid row side by side); the residual is the exact
 * deviation form of benchmark_2d.f90:465-618 (compute_update_exact).
 *
 * Threading: a handle owns device memory and one HIP stream; calls on one handle
 * must not overlap.  Host-array entry points block until the result is on the host.
 */
#ifndef FVWB_H
#define FVWB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define FVWB_OK               0
#define FVWB_ERR_ARG         (-1)  /* bad argument / unsupported combination   */
#define FVWB_ERR_HIP         (-2)  /* a HIP runtime call failed                */
#define FVWB_ERR_OOM         (-3)  /* device allocation failed                 */
#define FVWB_ERR_NONPHYSICAL (-4)  /* rho<=0, p<=0 or a non-finite value seen  */
#define FVWB_ERR_STATE       (-5)  /* call order (e.g. advance before upload)  */

/* ---- scheme selectors ---------------------------------------------------- */
/* boundary: FROZEN_RING is the reference's (benchmark_2d.f90:611-614: dudt=0 on
 * the outer ring of cells); REFLECT mirrors the deviation with the normal
 * momentum negated (solid walls). */
#define FVWB_BC_FROZEN_RING   0
#define FVWB_BC_REFLECT       1
/* reconstruction of the deviation at faces: FIRST_ORDER is the reference's
 * piecewise-constant state (benchmark_2d.f90:533-537); MUSCL_MINMOD uses the
 * minmod slope of limiters.f90:1-21 on conservative deviations. */
#define FVWB_RECON_FIRST_ORDER   0
#define FVWB_RECON_MUSCL_MINMOD  1
/* numerical flux: RUSANOV is compute_llflux (benchmark_2d.f90:353-367, wave
 * speed |v|+c with |v| the velocity magnitude); HLL is the two-wave HLL flux with
 * Davis speed bounds on the signed normal velocity; HLLC (DG only) is the
 * three-wave riemann_hllc of dg_with_source.f90:1318-1374 carried to 2D, the
 * tangential velocity advected by the contact. */
#define FVWB_FLUX_RUSANOV     0
#define FVWB_FLUX_HLL         1
#define FVWB_FLUX_HLLC        2
/* time integrator: RK2_HEUN is the reference's evolve (benchmark_2d.f90:246-250);
 * SSPRK3 is the Shu-Osher scheme of dg_with_source.f90:194-204. */
#define FVWB_RK2_HEUN         0
#define FVWB_SSPRK3           1
/* storage / arithmetic precision of the deviation state */
#define FVWB_FP64             0   /* plain FP64: full-flux differences, drop-in   */
#define FVWB_FP32             1   /* FP32 bulk: cancellation-free deviation flux  */

typedef struct fvwb_params {
  int32_t nvar;        /* must be 4                                           */
  int32_t nx, ny;      /* cells; nx, ny >= 4                                  */
  int32_t bc;          /* FVWB_BC_*                                           */
  int32_t recon;       /* FVWB_RECON_*                                        */
  int32_t flux;        /* FVWB_FLUX_*                                         */
  int32_t integrator;  /* FVWB_RK2_HEUN / FVWB_SSPRK3                         */
  int32_t precision;   /* FVWB_FP64 / FVWB_FP32                               */
  double  gamma;       /* ratio of specific heats (reference: 1.4 as an FP32
                          literal, i.e. 1.39999997615814209)                  */
  double  boxlen_x, boxlen_y;
  double  cfl;         /* dt = 0.5*h/cmax*cfl, h = min(dx, dy) (the reference's
                          :242 takes dx, the same on its square grids)         */
  double  rho0, p0, g; /* equilibrium constants; nequilibrium=1 is (1,1,1),
                          nequilibrium=2/3 is (1.21 as FP32 literal, 1, 1)     */
} fvwb_params;

typedef struct fvwb_ctx* fvwb_handle;

/* Fill `p` with the reference's benchmark_2d set-up (parameters_2d.f90:1-24):
 * nx x ny cells, gamma/rho0 rounded through FP32 as the reference's literals
 * are, first-order Rusanov, frozen ring, Heun RK2, FP64, cfl 0.5. */
int fvwb_params_reference_2d(fvwb_params* p, int32_t nx, int32_t ny, int32_t nequilibrium);

/* Create / destroy a single-instance context (device buffers + one stream). */
int fvwb_create(fvwb_handle* h, const fvwb_params* p);
int fvwb_destroy(fvwb_handle h);
/* Message of the last failure on `h` (or of the last failed create if h==NULL). */
const char* fvwb_last_error(fvwb_handle h);

/* ---- per-stage drop-in (host arrays, blocking) ---------------------------
 * Replaces `subroutine compute_update_exact(u, w_eq, dudt)`
 * (benchmark_2d.f90:465-618).  u: conservative state AoS(4,nx,ny); w_eq:
 * PRIMITIVE equilibrium at cell centres AoS(4,nx,ny), at rest; dudt (out):
 * AoS(4,nx,ny).  Uses the handle's recon/flux/bc (the reference set-up is
 * FIRST_ORDER/RUSANOV/FROZEN_RING).  Always computed in FP64.  The face
 * equilibrium sits where the reference puts it: x faces at (i-1)*dx, y faces at
 * (j-1)*dx (benchmark_2d.f90:507-514; the same as (j-1)*dy when dx == dy). */
int fvwb_compute_update_exact(fvwb_handle h, const double* u, const double* w_eq, double* dudt);

/* The residual L(u) that fvwb_advance integrates: as fvwb_compute_update_exact
 * but with the y faces at (j-1)*dy.  No separate counterpart in the reference
 * (identical to compute_update_exact when dx == dy, as in every shipped
 * set-up).  With dx != dy the reference's face positions put the equilibrium
 * of a y face at the wrong height; rest stays exact either way, since the same
 * positions enter the flux and the equilibrium flux, but a perturbation is
 * carried on the wrong background.  The resident loop uses the corrected
 * positions; a caller who needs the reference's evolve at dx != dy bit for bit
 * keeps its own loop over the drop-ins. */
int fvwb_compute_residual(fvwb_handle h, const double* u, const double* w_eq, double* dudt);

/* Replaces `subroutine compute_update(u, w_eq, dudt)` (benchmark_2d.f90:370-463),
 * the plain non-well-balanced residual: Rusanov between the cell states at every
 * face, the pointwise source with grad phi = (1,1), a frozen ring; FP64, the
 * reference's result bit for bit.  w_eq is accepted and unused, as there. */
int fvwb_compute_update(fvwb_handle h, const double* u, const double* w_eq, double* dudt);

/* Replaces `subroutine compute_max_speed(u, cmax)` (benchmark_2d.f90:264-279):
 * cmax = max over cells of sqrt(vx^2+vy^2) + sqrt(gamma*max(p,1e-10)/max(rho,1e-10)). */
int fvwb_compute_max_speed(fvwb_handle h, const double* u, double* cmax);

/* ---- device-resident stepping (replaces the loop of `evolve`,
 *      benchmark_2d.f90:221-260) --------------------------------------------
 * upload: dU = u - conservative(w_eq) is formed on the device in FP64 and
 * stored in the handle's precision; t is reset to 0.  The face equilibria and
 * the centre equilibrium entering the wave speed for dt are the closed form of
 * the params (as in every reference set-up, where w_eq is that closed form);
 * w_eq itself is kept only to form dU and to rebuild u on download, so a
 * state equal to conservative(w_eq) stays exactly at rest. */
int fvwb_upload(fvwb_handle h, const double* u, const double* w_eq);
/* Initialise on the device (no host arrays): dU = 0 except a pressure bump
 * dp = amp*exp(-width*((x-x0)^2+(y-y0)^2)), i.e. dE = dp/(gamma-1).  w_eq is
 * the closed-form equilibrium at cell centres. */
int fvwb_init_gaussian(fvwb_handle h, double amp, double x0, double y0, double width);
/* Advance at most `nsteps` steps while t < t_end (pass t_end = +inf for a fixed
 * count).  dt is computed on the device each step from the fused CFL reduction,
 * so the host does not synchronise between steps.  t_out / steps_done may be
 * NULL.  Returns FVWB_ERR_NONPHYSICAL if any step met rho<=0, p<=0 or NaN. */
int fvwb_advance(fvwb_handle h, int32_t nsteps, double t_end, double* t_out, int32_t* steps_done);
/* Same, but enqueue only (no host synchronisation, no status check); used for
 * timing with events on the handle's stream. */
int fvwb_advance_async(fvwb_handle h, int32_t nsteps, double t_end);
/* Copy back u = conservative(w_eq) + dU as AoS(4,nx,ny) FP64. */
int fvwb_download(fvwb_handle h, double* u);
/* Copy back the stored deviation dU as AoS(4,nx,ny) FP64 (exact widening). */
int fvwb_download_deviation(fvwb_handle h, double* du);
/* Simulated time and last dt after the enqueued steps (synchronises). */
int fvwb_get_time(fvwb_handle h, double* t, double* dt_last, int32_t* steps_done);
/* The handle's stream (a hipStream_t), for event timing by the caller. */
int fvwb_get_stream(fvwb_handle h, void** stream);

/* ---- launch shape (no counterpart in the reference) -------------------------
 * A wave owns a strip of 62*columns_per_lane columns and rows_per_strip rows.
 * columns_per_lane: 2 or 4 for FP32 (8- or 16-byte loads per lane and variable),
 * always 2 for FP64.  rows_per_strip: multiple of 4 in [4, 1024].  Results do not
 * depend on the shape; speed does. */
int fvwb_set_launch(fvwb_handle h, int32_t rows_per_strip, int32_t columns_per_lane);
int fvwb_get_launch(fvwb_handle h, int32_t* rows_per_strip, int32_t* columns_per_lane);

/* ---- measurement helper (no counterpart in the reference) ------------------
 * Device-to-device streaming copy of `bytes` (rounded down to 16), repeated
 * `reps` times on the null stream; *gbps = 2*bytes*reps / elapsed.  The
 * calibration point for achieved bandwidth and HBM counters.  Allocates and
 * frees its own two buffers. */
int fvwb_copy_bandwidth(uint64_t bytes, int32_t reps, double* gbps);


/* ---- batched instances (config "batch of independent problems") -------------
 * Many independent 2D problems of different sizes advance together: each has
 * its own nx, ny, box, rho0, p0, g, its own CFL time step and clock; all share
 * bc, recon, flux, integrator, precision, gamma and cfl (checked at create).
 * One persistent launch per Runge-Kutta stage covers every strip of every
 * instance; step pairs are replayed from a HIP graph.  No counterpart in the
 * reference (it runs one problem per executable); each instance computes what
 * fvwb_create/fvwb_advance would compute for it alone. */
typedef struct fvwb_batch* fvwb_batch_handle;
int fvwb_batch_create(fvwb_batch_handle* h, int32_t ninst, const fvwb_params* per_inst);
int fvwb_batch_destroy(fvwb_batch_handle h);
const char* fvwb_batch_last_error(fvwb_batch_handle h);
/* Device initial condition for every instance: closed-form equilibrium plus a
 * Gaussian pressure bump of amplitude amp[b] (as fvwb_init_gaussian); resets
 * every clock. */
int fvwb_batch_init_gaussian(fvwb_batch_handle h, const double* amp, double x0, double y0, double width);
/* Host state of one instance (as fvwb_upload); clocks restart at the next
 * fvwb_batch_reset_time. */
int fvwb_batch_upload(fvwb_batch_handle h, int32_t inst, const double* u, const double* w_eq);
int fvwb_batch_reset_time(fvwb_batch_handle h);
int fvwb_batch_advance(fvwb_batch_handle h, int32_t nsteps, double t_end);
int fvwb_batch_advance_async(fvwb_batch_handle h, int32_t nsteps, double t_end);
int fvwb_batch_download(fvwb_batch_handle h, int32_t inst, double* u);
int fvwb_batch_download_deviation(fvwb_batch_handle h, int32_t inst, double* du);
int fvwb_batch_get_time(fvwb_batch_handle h, int32_t inst, double* t, double* dt_last, int32_t* steps_done);
int fvwb_batch_get_stream(fvwb_batch_handle h, void** stream);
/* Replay step pairs from a HIP graph (default on). */
int fvwb_batch_use_graph(fvwb_batch_handle h, int32_t enable);
/* Launch shape: rows per strip (multiple of 4), FP32 columns per lane (2/4),
 * persistent blocks (0 = one per four strips). */
int fvwb_batch_set_launch(fvwb_batch_handle h, int32_t rows, int32_t cpl, int32_t blocks);
int fvwb_batch_get_launch(fvwb_batch_handle h, int32_t* rows, int32_t* cpl, int32_t* blocks, int32_t* ntiles);
/* Time `trial_steps` steps for each candidate shape on a saved copy of the
 * state (restored afterwards), keep the fastest and append it to the text file
 * cache_path (NULL: no cache); a cached entry for the same batch is reused. */
int fvwb_batch_autotune(fvwb_batch_handle h, int32_t trial_steps, const char* cache_path, int32_t* rows,
                        int32_t* cpl);

/* ---- 1D: benchmark_1d.f90 ------------------------------------------------
 * 1D Euler with gravity, gamma-law gas, potential phi = g*x on [0,boxlen],
 * state u(1:3,1:nx) = (rho, rho*v, E) with the variable index fastest.
 * Equilibria (nequilibrium, benchmark_1d.f90:134-155 for g = 1):
 *   1, 2  isothermal  rho = p = exp(-g*x), v = 0
 *   3     polytropic  rho = (1 - (gamma-1)/gamma*g*x)^(1/(gamma-1)), p = rho^gamma
 * A handle holds `ninst` independent instances that share recon, integrator and
 * bc; each has its own nx, boxlen, g, cfl and clock. */
#define FVWB1D_BC_PERIODIC       1  /* benchmark_1d.f90:318-321 codes; in the    */
#define FVWB1D_BC_ZERO_GRADIENT  2  /* reference they only change the two        */
#define FVWB1D_BC_REFLEXIVE      3  /* boundary faces (see fvwb1d_compute_update) */
#define FVWB1D_BC_WALL           4  /* configured scheme: solid walls, ghost cells
                                       mirror the deviation, momentum negated     */

typedef struct fvwb1d_params {
  int32_t nvar;          /* must be 3                                         */
  int32_t nx;            /* cells, nx >= 4 (fvwb1d_advance: nx <= 2048)       */
  int32_t bc;            /* FVWB1D_BC_*; the configured scheme takes
                            ZERO_GRADIENT or WALL                              */
  int32_t recon;         /* FVWB_RECON_FIRST_ORDER / FVWB_RECON_MUSCL_MINMOD  */
  int32_t integrator;    /* FVWB_RK2_HEUN (SSP-RK2) / FVWB_SSPRK3              */
  int32_t nequilibrium;  /* 1, 2 or 3 (above)                                  */
  double  gamma;         /* reference: 1.4 as an FP32 literal                  */
  double  boxlen;        /* reference: 1                                       */
  double  cfl;           /* dt = cfl*dx/max(|v|+c), last step clipped to t_end */
  double  g;             /* reference: 1                                       */
} fvwb1d_params;

typedef struct fvwb1d* fvwb1d_handle;

/* Fill `p` with benchmark_1d.f90's set-up (parameters.f90:1-22): nx cells,
 * gamma = (double)1.4f, boxlen 1, g 1, zero-gradient bc code, first order,
 * SSP-RK2, cfl 0.5. */
int fvwb1d_params_reference(fvwb1d_params* p, int32_t nx, int32_t nequilibrium);
int fvwb1d_create(fvwb1d_handle* h, int32_t ninst, const fvwb1d_params* per_inst);
int fvwb1d_destroy(fvwb1d_handle h);
const char* fvwb1d_last_error(fvwb1d_handle h);

/* Replaces `subroutine compute_update(u, w_eq, dudt)` of benchmark_1d.f90:263-377
 * (the EQL exact-deviation residual) for instance 0: u conservative AoS(3,nx),
 * w_eq PRIMITIVE equilibrium at cell centres AoS(3,nx), dudt (out) AoS(3,nx).
 * First order, Rusanov flux (compute_llflux :438-451), face states U_eq(face) +
 * (u - U_eq(centre)), source -rho*(x(i+1)-x(i-1))/(2/nx) (get_source :379-405,
 * evaluated on the conservative state as the reference does), and
 * dudt(1) = dudt(2), dudt(nx) = dudt(nx-1) (:374-375); the bc code changes only
 * the two boundary faces, whose fluxes reach no other cell, so it does not
 * change dudt.  Always FP64, rounding statement by statement as the reference. */
int fvwb1d_compute_update(fvwb1d_handle h, const double* u, const double* w_eq, double* dudt);
/* Replaces `subroutine compute_update_sr(u, w_eq, dudt)` (benchmark_1d.f90:553-747),
 * the solver 'WB1' that the reference selects by default (parameters.f90:8), for
 * instance 0 (w_eq is unused there, so it is not an argument here).  Hydrostatic
 * reconstruction: each cell's face states keep its entropy max(p,1e-5)/max(rho,
 * 1e-5)^gamma and its enthalpy plus phi = x (g = 1, as written in the reference's
 * phi), with the cell's velocity; Rusanov at the interior faces; source
 * ((p_R - p_L)*nx, -rho*v*(x(i+1)-x(i-1))*nx/2) on momentum and energy
 * (get_source_rg :409-434); dudt(1) = dudt(2), dudt(nx) = dudt(nx-1).  Uses the
 * handle's nx, gamma and boxlen only.  FP64 in the reference's statement order;
 * the device pow makes it agree to a few ulps rather than bit for bit.  An
 * isentropic rest state rho = (1-(gamma-1)/gamma*x)^(1/(gamma-1)), p = rho^gamma
 * gives dudt at rounding level. */
int fvwb1d_compute_update_sr(fvwb1d_handle h, const double* u, double* dudt);
/* Replaces `subroutine compute_update_fvm(u, w_eq, dudt)` (benchmark_1d.f90:454-549),
 * the plain solver 'FVM', for instance 0 (w_eq is unused there): Rusanov between
 * the cell states, source -rho*(x(i+1)-x(i-1))*nx/2 on momentum and the same
 * times the momentum on energy (get_source :379-405 called on the conservative
 * state), dudt(1) = dudt(2), dudt(nx) = dudt(nx-1).  Bit for bit. */
int fvwb1d_compute_update_fvm(fvwb1d_handle h, const double* u, double* dudt);
/* Replaces `subroutine compute_update(u, dudt)` of the stand-alone program fvm.f90
 * (:188-264), for instance 0: Rusanov between the cell states with the handle's
 * bc, FVWB1D_BC_PERIODIC or FVWB1D_BC_ZERO_GRADIENT (:200-230; other codes give
 * FVWB_ERR_ARG, the reference indexes out of bounds there); `source` is
 * fvm_commons' switch, 1 none, 2 gravity: -rho*(x+ - x-)/(2 dx) on momentum and
 * -rho*v*(x+ - x-)/(2 dx) on energy, x- and x+ clamped to the cell's centre at the
 * two ends (:229-236, compute_source :253-264).  Bit for bit. */
int fvwb1d_fvm_compute_update(fvwb1d_handle h, int32_t source, const double* u, double* dudt);
/* Replaces `subroutine compute_max_speed(u, cmax)` (benchmark_1d.f90:157-184,
 * and the identical fvm.f90:320-336) for instance 0: cmax = max |v| + sqrt(gamma*max(p,1e-10)/max(rho,1e-10)). */
int fvwb1d_compute_max_speed(fvwb1d_handle h, const double* u, double* cmax);
/* Residual of the configured scheme for instance `inst` (what fvwb1d_advance
 * integrates): deviation form with face equilibrium, p.recon, Rusanov, source
 * (0, -g*drho, -g*dm), ghost cells per p.bc.  No counterpart in the reference. */
int fvwb1d_compute_residual(fvwb1d_handle h, int32_t inst, const double* u, const double* w_eq, double* dudt);

/* Device-resident stepping.  upload: host state of one instance (w_eq as
 * above); init_gaussian: equilibrium plus a pressure pulse
 * amp[b]*exp(-(x-x0)^2/(2*width^2)) in every instance b (fvm.f90:124
 * convention); clocks restart at 0. */
int fvwb1d_upload(fvwb1d_handle h, int32_t inst, const double* u, const double* w_eq);
int fvwb1d_init_gaussian(fvwb1d_handle h, const double* amp, double x0, double width);
/* Advance every instance until its t reaches t_end (the last step clipped) or
 * it has taken nsteps more steps: one launch, one workgroup per instance, the
 * whole run in LDS.  Blocks until done; FVWB_ERR_NONPHYSICAL if any instance
 * produced rho<=0, p<=0 or a non-finite value (its clock stops there). */
int fvwb1d_advance(fvwb1d_handle h, int32_t nsteps, double t_end);
int fvwb1d_download(fvwb1d_handle h, int32_t inst, double* u);
int fvwb1d_get_time(fvwb1d_handle h, int32_t inst, double* t, double* dt_last, int32_t* steps_done);
int fvwb1d_get_stream(fvwb1d_handle h, void** stream);

/* ---- 2D modal DG in deviation form (config: P2, 3x3 Gauss-Legendre) -------
 * Modal discontinuous Galerkin for the same 2D Euler-with-gravity problem as
 * above (phi = g*(x+y), isothermal atmosphere).  The unknown is the deviation
 * dU = U - U_eq expanded in the orthonormal Legendre basis of
 * 2d/legendre.f90:1-50 (phi_n = sqrt(2n+1) P_n, so 1/4 of the integral of
 * phi_i(xi)phi_j(eta)phi_k(xi)phi_l(eta) over [-1,1]^2 is delta_ik delta_jl),
 * total degree <= k: modes (0,0) (1,0) (0,1) (2,0) (1,1) (0,2) in that order.
 * Volume and source integrals on (k+1)^2 Gauss-Legendre points, face integrals
 * on k+1 points per face (2d/benchmark_2d_dg.f90:1207-1244, 1372-1411,
 * 1426-1462: dudt = (vol_x/dx + vol_y/dy - edges)/2 + src/4); the fluxes and
 * source are deviations from the equilibrium's (dg_with_source.f90:1749-2031),
 * so a resting atmosphere is an exact fixed point.  Host arrays of modes are
 * the reference's layout delta_u(1:4,1:nx,1:ny,1:nmodes): variable fastest,
 * mode index slowest. */
#define FVWB_DG_LIMITER_NONE  0
#define FVWB_DG_LIMITER_TVB   1   /* 2d/limiters.f90:30-54 generalized_minmod on the
                                     linear modes of each conservative deviation;
                                     a limited cell drops its quadratic modes   */
#define FVWB_DG_LIMITER_POSITIVITY 2   /* compute_positivity, 2d/limiters.f90:478-654
                                     (limiter_type 'ONP', the reference's shipped
                                     limiter) on the full state U_eq + dU: on the
                                     set {Gauss-Lobatto x Gauss} u {Gauss x
                                     Gauss-Lobatto} u {the volume Gauss points},
                                     the non-mean density modes are scaled by
                                     min(|(rho_avg - eps_r)/(rho_avg - rho_min)|,
                                     1), then every non-mean mode by the smallest
                                     root of solve_for_t (:312-362) over the points
                                     whose pressure is not above eps_p; eps_r and
                                     eps_p are 1e-5 of the cell's equilibrium
                                     density and pressure (the reference's 1e-10 is
                                     below what an FP32 deviation resolves); the
                                     equilibrium's own modes are its Gauss
                                     projection; means never change            */
#define FVWB_DG_LIMITER_TVB_POSITIVITY 3   /* TVB, then positivity              */
#define FVWB_DG_LIMITER_CHARACTERISTIC 4   /* the TVB minmod of the linear modes
                                     taken in characteristic variables: the
                                     x slope and the x differences of the
                                     neighbour means (of the deviation) through
                                     the left eigenvectors of dF/dU at the cell's
                                     full-state mean, the y ones through those of
                                     dG/dU, limited componentwise and mapped back
                                     with the right eigenvectors; a limited cell
                                     drops its quadratic modes (the per-direction
                                     form of the characteristic limiting the
                                     reference sketches in 2d/benchmark_2d_dg.f90:
                                     2021-2089 and 2d/limiters.f90:1200-1295)   */
#define FVWB_DG_LIMITER_CHARACTERISTIC_POSITIVITY 5   /* characteristic, then positivity */
#define FVWB_DG_SSPRK3        0   /* Shu-Osher SSP-RK3 (dg_with_source.f90:194-204)  */
#define FVWB_DG_SSPRK54       1   /* five-stage fourth-order SSP Runge-Kutta of
                                     2d/benchmark_2d_dg.f90:683-710 (solver 'RK4',
                                     the reference's shipped DG integrator)     */

typedef struct fvwb_dg_params {
  int32_t nvar;        /* must be 4                                             */
  int32_t nx, ny;      /* cells, >= 2                                           */
  int32_t degree;      /* polynomial degree k (the GPU path implements k = 2)   */
  int32_t flux;        /* FVWB_FLUX_RUSANOV / FVWB_FLUX_HLL / FVWB_FLUX_HLLC    */
  int32_t limiter;     /* FVWB_DG_LIMITER_*                                     */
  int32_t bc;          /* FVWB_BC_REFLECT (walls)                               */
  int32_t integrator;  /* FVWB_DG_SSPRK3 / FVWB_DG_SSPRK54                      */
  double  gamma, boxlen_x, boxlen_y;
  double  cfl;         /* dt = cfl / ((2k+1) max((|u|+c)/dx + (|v|+c)/dy))      */
  double  rho0, p0, g; /* equilibrium rho0*exp(-k(x+y)), k = rho0*g/p0          */
  double  tvb_m;       /* TVB constant M: slopes below M*dx^2 are kept          */
} fvwb_dg_params;

typedef struct fvwb_dg* fvwb_dg_handle;
/* The device path: FP32 modes, arithmetic and equilibrium tables, the
 * projections on the matrix cores, degree 2 only. */
int fvwb_dg_create(fvwb_dg_handle* h, const fvwb_dg_params* p);
int fvwb_dg_destroy(fvwb_dg_handle h);
const char* fvwb_dg_last_error(fvwb_dg_handle h);
/* dudt of the deviation modes (host AoS(4,nx,ny,6) FP64 in and out): the
 * counterpart of compute_update (2d/benchmark_2d_dg.f90:1137-1479) for the
 * deviation form, P2 basis and the face flux of p.flux. */
int fvwb_dg_compute_residual(fvwb_dg_handle h, const double* modes, double* dudt);
/* Device state: upload deviation modes (limited on arrival unless p.limiter is
 * FVWB_DG_LIMITER_NONE), or build them from a pressure perturbation
 * dp = amp * sum_q a[q] cos(2 pi (mk[q] x + nk[q] y) + theta[q]) / sum_q |a[q]|
 * (dE = dp/(gamma-1)) projected on P2; the clock restarts at 0. */
int fvwb_dg_upload(fvwb_dg_handle h, const double* modes);
int fvwb_dg_init_fourier(fvwb_dg_handle h, int32_t nterms, const double* a, const double* mk, const double* nk,
                         const double* theta, double amp);
/* nsteps steps of p.integrator (limiter after every stage; dt = cfl/(5 max rate)
 * from the device for either integrator, the last step clipped to t_end; no host
 * synchronisation inside). */
int fvwb_dg_advance(fvwb_dg_handle h, int32_t nsteps, double t_end);
int fvwb_dg_advance_async(fvwb_dg_handle h, int32_t nsteps, double t_end);
int fvwb_dg_download(fvwb_dg_handle h, double* modes);
int fvwb_dg_get_time(fvwb_dg_handle h, double* t, double* dt_last, int32_t* steps_done);
int fvwb_dg_get_stream(fvwb_dg_handle h, void** stream);

/* ---- 1D modal DG with source (dg_with_source.f90, dg_commons.f90) ---------
 * Modes delta_u(1:nvar,1:n,1:nx) (variable fastest, then mode, then cell) of
 * the deviation from the equilibrium, over the basis of legendre.f90:1-24
 * (phi_0 = sqrt(0.5), phi_1 = sqrt(6)/2 x, phi_2 = sqrt(10)/4 (3x^2-1) with the
 * square roots taken in FP32 as the reference writes them, phi_3.. the plain
 * Legendre polynomials), with n Gauss points from gl_quadrature
 * (legendre.f90:77-127).  The equilibrium is rho = p = exp(-x), v = 0
 * (get_eq_solution :1031-1048); u_eq(1:nvar,1:n,1:nx) holds it, conservative,
 * at each cell's Gauss points, as the reference's main program builds it. */
#define FVWB_DG1D_RIEMANN_LLF   1   /* riemann_llf  (dg_with_source.f90:1299-1316) */
#define FVWB_DG1D_RIEMANN_HLLC  2   /* riemann_hllc (dg_with_source.f90:1318-1374) */

typedef struct fvwb_dg1d_params {
  int32_t nvar;      /* must be 3                                              */
  int32_t nx;        /* cells, >= 3                                            */
  int32_t n;         /* modes = Gauss points, 1..6 (dg_commons.f90:4-7)        */
  int32_t riemann;   /* FVWB_DG1D_RIEMANN_*                                    */
  int32_t source;    /* 1 none, 2 gravity phi = x (compute_source :1196-1211)  */
  int32_t pad;
  double  gamma, boxlen;
} fvwb_dg1d_params;

typedef struct fvwb_dg1d* fvwb_dg1d_handle;
/* dg_commons.f90 set-up: n = 3, HLLC, source 2, gamma = 1.4 and boxlen = 1
 * rounded as the reference's default-real literals. */
int fvwb_dg1d_params_reference(fvwb_dg1d_params* p, int32_t nx);
int fvwb_dg1d_create(fvwb_dg1d_handle* h, const fvwb_dg1d_params* p);
int fvwb_dg1d_destroy(fvwb_dg1d_handle h);
const char* fvwb_dg1d_last_error(fvwb_dg1d_handle h);
/* Replaces `subroutine compute_update_exact_delta(delta_u, u_eq, dudt)`
 * (dg_with_source.f90:1749-2031), the well-balanced deviation-form DG residual:
 * volume terms F(u_eq + delta) - F(u_eq) at the Gauss points, face states
 * u_eq(face) + delta trace, numerical flux minus the equilibrium's physical
 * flux at the faces, source s(u_eq + delta) - s(u_eq); dudt of the two end
 * cells is zero (:2027-2028), so neither the boundary faces nor the bc code
 * reach it.  The reference's scaling is kept: its basis is orthonormal on
 * [-1,1] and its rate is (1/dx)(volume - faces) + (1/2) source, half the usual
 * DG rate.  FP64, the reference's statement order: equal to it bit for bit
 * (the face equilibria and the tables are computed on the host with the
 * reference's expressions). */
int fvwb_dg1d_compute_update_exact_delta(fvwb_dg1d_handle h, const double* delta_u, const double* u_eq,
                                         double* dudt);
/* The main program's time loop with its shipped integrator 'RKi'
 * (dg_with_source.f90:191-199, 264-288, 303-334), resident on the device:
 * upload the modes, u_eq and u_init (the nodal state the first dt is taken
 * from: the main program's initial Gauss-point values); each step takes
 * dt = 0.9*dx/cmax/(2n+1) with cmax the largest |v|+c over the cells at the
 * first Gauss point of u_init, runs the five SSP(5,4) stages of
 * compute_update_exact_delta with the reference's FP32-rounded coefficients,
 * then sets u_init = u_eq + the modes evaluated, as the reference does, at
 * legendre(x_quad) with x_quad the point's physical coordinate.  advance runs
 * while t < t_end, at most nsteps more steps; it returns FVWB_ERR_NONPHYSICAL
 * and stops the clock if cmax or dt stops being finite and positive.  Equal to
 * the reference's run bit for bit (every t and dt). */
int fvwb_dg1d_upload(fvwb_dg1d_handle h, const double* delta_u, const double* u_eq, const double* u_init);
int fvwb_dg1d_advance(fvwb_dg1d_handle h, int32_t nsteps, double t_end);
/* the modes and (when u_init is not NULL) the last u_init */
int fvwb_dg1d_download(fvwb_dg1d_handle h, double* delta_u, double* u_init);
int fvwb_dg1d_get_time(fvwb_dg1d_handle h, double* t, double* dt_last, int32_t* steps_done);

/* ---- the reference's own 2D DG program (2d/benchmark_2d_dg.f90) ----------------
 * Tensor-product (Q) Legendre basis with mx = my = order modes per direction,
 * FP64, the reference's layouts: modes and nodal values
 * delta_u(1:4,1:nx,1:ny,1:order,1:order) (element (v,i,j,a,b) at
 * v + 4*(i + nx*(j + ny*(a + order*b))), mode / node indices slowest), node
 * coordinates x, y(1:nx,1:ny,1:order,1:order) as get_coords (:93-120) builds
 * them.  The switches are the compile-time constants of 2d/parameters_dg_2d.f90.
 * The reference's residual indexes the x faces with the y bounds (:1338-1339)
 * and sizes get_source's work array nx x nx (:1563), so it is defined for
 * nx == ny only; create refuses other sizes.  Results equal the reference's
 * expressions bit for bit (the basis and Gauss rule are built on the host with
 * the reference's own expressions; the device does +, -, *, / and sqrt in the
 * reference's order, without contraction). */
#define FVWB_DGQ_FLUX_NONE 0   /* a flux_type matching no branch of compute_num_flux
                                  (:991-1006), e.g. the shipped 'llf': F = G = 0 */
#define FVWB_DGQ_FLUX_LLF1 1   /* 'llf1': compute_llflux, signed normal velocity + c */
#define FVWB_DGQ_FLUX_HLL2 2   /* 'hll2': compute_hllflux (:1008-1026)                */
#define FVWB_DGQ_FLUX_HLLC 3   /* 'hllc': compute_hllcflux (:1030-1134) as written   */
#define FVWB_DGQ_SOLVER_RK4 0  /* 'RK4' (shipped): five-stage SSP, :683-710          */
#define FVWB_DGQ_SOLVER_SS4 1  /* 'SS4': the same scheme, other literals, :711-735  */
#define FVWB_DGQ_SOLVER_EQL 2  /* 'EQL': Heun, :672-681                              */
#define FVWB_DGQ_SOLVER_DEB 3  /* 'DEB': one forward Euler stage, :737-747            */

typedef struct fvwb_dgq_params {
  int32_t nx, ny;          /* cells, nx == ny                                        */
  int32_t order;           /* mx = my: 2 or 3 (the reference does not compile with 1:
                              2d/limiters.f90 indexes the second mode)               */
  int32_t flux_type;       /* FVWB_DGQ_FLUX_*                                        */
  int32_t source;          /* 1: none; 2: get_source with grad_phi; 3: get_adv_source */
  int32_t grad_phi_case;   /* 1: grad_p = (x, y); 2: point mass at (3, 3)            */
  int32_t bc;              /* 1: periodic; 2, 3: transmissive (get_boundary_conditions) */
  int32_t ninit;           /* 12 freezes dudt outside r = 2 (special_boundary_conditions);
                              other values leave compute_update alone                */
  int32_t solver;          /* FVWB_DGQ_SOLVER_*                                      */
  int32_t use_limiter;     /* 1: apply_limiter with limiter_type 'ONP', i.e.
                              compute_positivity (2d/limiters.f90:478-654)           */
  double gamma, boxlen_x, boxlen_y, cfl;
  double eps;              /* compute_positivity's threshold                         */
} fvwb_dgq_params;

typedef struct fvwb_dgq* fvwb_dgq_handle;

/* The shipped set-up of 2d/parameters_dg_2d.f90 (nx = ny = 8, mx = my = 2,
 * 'llf', source 1, grad_phi_case 2, bc 1, ninit 1, 'RK4', 'ONP', gamma 1.4 and
 * eps 1e-10 as default-real literals, cfl 0.2, unit box). */
int fvwb_dgq_params_reference(fvwb_dgq_params* p);
/* 2d/legendre.f90:77-108 gl_quadrature(x_quad, w_quad, n), n = 1..6, on the
 * host with the reference's expressions (no device work): the nodes and
 * weights get_coords and every drop-in below use. */
int fvwb_dgq_gl_quadrature(int32_t n, double* x_quad, double* w_quad);
int fvwb_dgq_create(fvwb_dgq_handle* out, const fvwb_dgq_params* p);
int fvwb_dgq_destroy(fvwb_dgq_handle h);
const char* fvwb_dgq_last_error(fvwb_dgq_handle h);
/* 2d/benchmark_2d_dg.f90:1137-1479 compute_update(delta_u, x, y, u_eq, dudt);
 * u_eq is unused by the reference and may be NULL. */
int fvwb_dgq_compute_update(fvwb_dgq_handle h, const double* delta_u, const double* x, const double* y,
                            const double* u_eq, double* dudt);
/* :826-870 compute_max_speed(u, cs_max, v_xmax, v_ymax, speed_max) on
 * u(1:4,1:nx,1:ny), with the reference's bookkeeping: v_xmax, v_ymax from the
 * last cell (x outer, y inner) whose speed reaches the running maximum, cs_max
 * the smallest sound speed from that cell on. */
int fvwb_dgq_compute_max_speed(fvwb_dgq_handle h, const double* u, double* cs_max, double* v_xmax,
                               double* v_ymax, double* speed_max);
/* :1516-1555 apply_limiter(u) in place (compute_positivity when use_limiter) */
int fvwb_dgq_apply_limiter(fvwb_dgq_handle h, double* u);
/* :497-592 get_modes_from_nodes(nodes, u) and get_nodes_from_modes(u, nodes) */
int fvwb_dgq_modes_from_nodes(fvwb_dgq_handle h, const double* nodes, double* modes);
int fvwb_dgq_nodes_from_modes(fvwb_dgq_handle h, const double* modes, double* nodes);
/* :624-775 evolve, resident on the device: upload takes the limited modes and
 * the node coordinates and resets the clock; advance takes at most nsteps
 * steps while t < t_end, each with the dt of :671 from the device's
 * compute_max_speed of the means, the configured solver and apply_limiter
 * after every stage; FVWB_ERR_NONPHYSICAL stops the clock if dt stops being
 * finite and positive. */
int fvwb_dgq_upload(fvwb_dgq_handle h, const double* delta_u, const double* x, const double* y);
int fvwb_dgq_advance(fvwb_dgq_handle h, int32_t nsteps, double t_end);
int fvwb_dgq_download(fvwb_dgq_handle h, double* delta_u);
int fvwb_dgq_get_time(fvwb_dgq_handle h, double* t, double* dt_last, int64_t* steps_done);
/* :23-89 compute_error: per variable the maximum |u - u_init| (linf), the
 * Gauss-weighted L1 norm sum_cells (sum_ab |u - u_init| w_a w_b)(dx dy)/4 and
 * the L2 norm, with dx = 1/nx, dy = 1/ny as the reference takes them; a device
 * reduction (its summation order differs from the reference's serial loop, so
 * the norms agree to rounding).  nodes may be NULL: then the nodal values of
 * the resident state are used. */
int fvwb_dgq_compute_error(fvwb_dgq_handle h, const double* nodes, const double* u_init, double* linf,
                           double* l1, double* l2);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops): compute_error of
 * device arrays with no handle and no allocation.  nodes and u_init are device
 * arrays (4, nx, ny, order, order) as above, out a device array of 12
 * doubles (linf[4], l1[4], l2[4]), equal bit for bit to fvwb_dgq_compute_error;
 * stream is a hipStream_t (NULL: the default stream) and the call returns
 * once the work is queued on it.  The block partial sums live in a static
 * device array of 384 KiB per GPU, so nx*ny is limited to
 * FVWB_DGQ_NOALLOC_MAX_CELLS (FVWB_ERR_ARG beyond it); calls that overlap on
 * one GPU share that array and race. */
#define FVWB_DGQ_NOALLOC_MAX_CELLS (1 << 20)
int fvwb_dgq_compute_error_noalloc(int32_t nx, int32_t ny, int32_t order, const double* nodes, const double* u_init,
                                   double* out, void* stream);

/* ---- checkpoint / resume ----------------------------------------------------
 * No counterpart in the reference, whose snapshots are six-digit ASCII and not
 * restartable (benchmark_2d.f90:115-143).  save writes the device state a time
 * loop carries between steps (the state at t^n, the equilibrium tables loaded
 * with it, the clock and the step parity) to a binary file; load restores it
 * into a handle created with the same parameters (checked byte for byte;
 * FVWB_ERR_ARG otherwise, or for a missing, truncated or foreign file).  A
 * loaded handle continues exactly as the saved one would have: the same
 * steps, bit for bit.  save needs uploaded state (FVWB_ERR_STATE otherwise);
 * a failed load leaves the handle without state. */
int fvwb_checkpoint_save(fvwb_handle h, const char* path);
int fvwb_checkpoint_load(fvwb_handle h, const char* path);
int fvwb1d_checkpoint_save(fvwb1d_handle h, const char* path);
int fvwb1d_checkpoint_load(fvwb1d_handle h, const char* path);
int fvwb_dg_checkpoint_save(fvwb_dg_handle h, const char* path);
int fvwb_dg_checkpoint_load(fvwb_dg_handle h, const char* path);
int fvwb_dg1d_checkpoint_save(fvwb_dg1d_handle h, const char* path);
int fvwb_dg1d_checkpoint_load(fvwb_dg1d_handle h, const char* path);

/* ---- 2D residual distribution (benchmark_rd.f90, solver 'RD1') ---------------
 * A vertex scheme on nx x ny nodes x_i = (i-1) boxlen_x/(nx-1), y_j likewise,
 * state u(1:4, 1:nx, 1:ny) as the reference holds it with mx = my = 1 (AoS,
 * FP64, x fastest after the variable).  The parameters are those the
 * reference reads from its module parameters_rd, which it does not ship. */
#define FVWB_RD_BC_PERIODIC     2   /* benchmark_rd.f90:723-727: copies, x then y */
#define FVWB_RD_BC_TRANSMISSIVE 3   /* benchmark_rd.f90:729-733 */

typedef struct {
  int32_t nx, ny;      /* nodes per side, >= 2 */
  int32_t bc;          /* FVWB_RD_BC_PERIODIC or FVWB_RD_BC_TRANSMISSIVE (others stop the reference) */
  int32_t pad;
  double gamma, boxlen_x, boxlen_y, cfl;
} fvwb_rd_params;

typedef struct fvwb_rd* fvwb_rd_handle;

int fvwb_rd_create(fvwb_rd_handle* out, const fvwb_rd_params* p);
int fvwb_rd_destroy(fvwb_rd_handle h);
const char* fvwb_rd_last_error(fvwb_rd_handle h);
/* benchmark_rd.f90:612-745 compute_update_rd(u, diff, dt, x, y, dudt): host
 * arrays AoS(4,nx,ny); the reference's x and y only feed get_source's unused
 * coordinates.  dudt equals the reference's expressions bit for bit. */
int fvwb_rd_compute_update(fvwb_rd_handle h, const double* u, const double* diff, double dt, double* dudt);
/* benchmark_rd.f90:448-475 compute_max_speed(u, speed_max) */
int fvwb_rd_compute_max_speed(fvwb_rd_handle h, const double* u, double* cmax);
/* benchmark_rd.f90:333-388 evolve with solver 'RD1', state resident on the
 * device: upload resets the clock; advance takes at most nsteps steps while
 * t < t_end (dt = min(cfl dx dy/cmax/3, t_end - t), two update passes per
 * step); FVWB_ERR_NONPHYSICAL if the wave-speed maximum is not finite and
 * positive. */
int fvwb_rd_upload(fvwb_rd_handle h, const double* u);
int fvwb_rd_advance(fvwb_rd_handle h, int32_t nsteps, double t_end);
int fvwb_rd_download(fvwb_rd_handle h, double* u);
int fvwb_rd_get_time(fvwb_rd_handle h, double* t, double* dt_last, int64_t* steps_done);
/* The checkpoint format above for the resident state and clock of this
 * handle; no counterpart in the reference, whose output is ASCII snapshots. */
int fvwb_rd_checkpoint_save(fvwb_rd_handle h, const char* path);
int fvwb_rd_checkpoint_load(fvwb_rd_handle h, const char* path);

#ifdef __cplusplus
}
#endif
#endif /* FVWB_H */
