/*
 * fvwb_oracle.h — CPU FP64 oracle for the well-balanced Euler-with-gravity path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product (libfvwb.so) never links it.
 *
 * Arrays use the reference's Fortran layout: u(1:nvar,1:nx,1:ny) with the
 * variable index fastest.  Two families:
 *   orc_ref2d_*  — statement-by-statement restatements of benchmark_2d.f90
 *                  (same operation order, no FMA contraction), pinned against the
 *                  flang build of the reference in oracle/_ref;
 *   orc_fv2d_*   — the configured scheme (MUSCL-minmod / HLL / reflecting walls /
 *                  SSP-RK3) in the same plain FP64 form, pinned by invariants
 *                  and closed forms in tests/;
 *   orc_ref1d_*  — restatements of benchmark_1d.f90 (its 'EQL', 'WB1' and 'FVM'
 *                  residuals), pinned like orc_ref2d_*;
 *   orc_fvm_*    — the stand-alone 1D program fvm.f90, pinned likewise;
 *   orc_fv1d_*   — the configured 1D scheme (deviation form, MUSCL-minmod,
 *                  Rusanov, SSP-RK2/RK3, walls), pinned by invariants;
 *   orc_dg_*, orc_ref_dg2d_* — the reference's Legendre basis, Gauss-Legendre
 *                  quadrature and tensor projections (2d/), pinned like above;
 *   orc_dg2d_*   — the configured modal DG (deviation form, P_k, k <= 2),
 *                  pinned by invariants;
 *   orc_dg1d_*   — restatements of the 1D modal DG of dg_with_source.f90 and
 *                  its legendre.f90, pinned like orc_ref2d_*.
 */
#ifndef FVWB_ORACLE_H
#define FVWB_ORACLE_H

#include "../include/fvwb.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- pointwise physics (benchmark_2d.f90:145-171, 283-367) -------------- */
void   orc_ref_primitive(double gamma, const double u[4], double w[4]);
void   orc_ref_conservative(double gamma, const double w[4], double u[4]);
double orc_ref_speed(double gamma, const double u[4]);
void   orc_ref_flux(double gamma, const double u[4], double fx[4], double fy[4]);
void   orc_ref_llflux(double gamma, const double ul[4], const double ur[4],
                      const double fl[4], const double fr[4], double out[4]);
/* limiters.f90:1-21 */
double orc_minmod(double x, double y, double z);
/* 2d/limiters.f90:30-54 (TVB) with M and dx given explicitly */
double orc_generalized_minmod(double x, double y, double z, double M, double dx);
/* 2d/benchmark_2d_dg.f90:1008-1026 compute_hllflux, reproduced as written
 * (wave-speed bounds from the velocity magnitude); flag 1 = x, 2 = y. */
void   orc_ref_hllflux(double gamma, const double ul[4], const double ur[4],
                       const double fl[4], const double fr[4], double out[4]);

/* dg_with_source.f90:1318-1374 riemann_hllc (1D, u = (rho, rho v, E)), as written */
void   orc_ref_riemann_hllc_1d(double gamma, const double ul[3], const double ur[3], double f[3]);
/* its 2D extension used by the configured schemes (flux FVWB_FLUX_HLLC): the same
 * wave speeds and acoustic star state along direction dir (0 x, 1 y), the
 * tangential velocity taken from the upwind side of the contact */
void   orc_hllc_2d(double gamma, int dir, const double ul[4], const double ur[4], double f[4]);

/* ---- benchmark_2d.f90 restatements --------------------------------------- */
/* get_equilibrium_solution (174-218) at nx*ny points (x,y given per point). */
void   orc_ref2d_equilibrium(int nequilibrium, int n, const double* x, const double* y, double* w);
/* compute_update_exact (465-618) on the reference's set-up: nequilibrium
 * selects the closed form of the face equilibrium; gamma and box from p. */
/* :370-463 compute_update, the plain non-well-balanced residual */
void   orc_ref2d_update_plain(const fvwb_params* p, const double* u, double* dudt);
void   orc_ref2d_update_exact(const fvwb_params* p, int nequilibrium,
                              const double* u, const double* w_eq, double* dudt);
/* compute_max_speed (264-279) */
double orc_ref2d_max_speed(const fvwb_params* p, const double* u);
/* evolve (221-260): Heun RK2 while t < t_end, at most max_steps steps. */
void   orc_ref2d_evolve(const fvwb_params* p, int nequilibrium, double* u, const double* w_eq,
                        double t_end, int max_steps, double* t_out, int* steps_out);

/* ---- configured scheme (plain FP64 form) --------------------------------- */
/* closed-form equilibrium of p (rho0,p0,g) at one point, primitive */
void   orc_fv2d_equilibrium(const fvwb_params* p, double x, double y, double w[4]);
/* residual L(u) with p's recon / flux / bc; w_eq primitive at centres */
void   orc_fv2d_residual(const fvwb_params* p, const double* u, const double* w_eq, double* dudt);
/* numerical flux minus equilibrium flux at one face (dir 0 = x, 1 = y) */
void   orc_fv2d_face_flux(const fvwb_params* p, int dir, const double ueq[4], const double dl[4],
                          const double dr[4], double out[4]);
/* one step of p's integrator with the given dt */
void   orc_fv2d_step(const fvwb_params* p, double* u, const double* w_eq, double dt);
/* max over cells of |v|+c (the reference's compute_max_speed for any gamma) */
double orc_fv2d_max_speed(const fvwb_params* p, const double* u);
/* advance: dt = 0.5*min(dx,dy)/cmax*cfl each step, integrator from p.
 * Runs while t < t_end and steps < max_steps. */
void   orc_fv2d_advance(const fvwb_params* p, double* u, const double* w_eq,
                        double t_end, int max_steps, double* t_out, int* steps_out,
                        double* dt_last);

/* ---- 1D: benchmark_1d.f90 restatements (nvar = 3, AoS(3,nx)) ------------ */
void   orc_ref1d_primitive(double gamma, const double u[3], double w[3]);      /* :115-123 */
void   orc_ref1d_conservative(double gamma, const double w[3], double u[3]);   /* :125-132 */
double orc_ref1d_speed(double gamma, const double u[3]);                       /* :171-184 */
void   orc_ref1d_flux(double gamma, const double u[3], double f[3]);           /* :186-198 */
void   orc_ref1d_llflux(double gamma, const double ul[3], const double ur[3],   /* :438-451 */
                        const double fl[3], const double fr[3], double out[3]);
/* get_equilibrium_solution (:134-155) at n points */
void   orc_ref1d_equilibrium(double gamma, int nequilibrium, int n, const double* x, double* w);
/* compute_update, solver 'EQL' (:263-377), with p->nx, gamma, boxlen, bc code
 * (1 periodic, 2 zero gradient, 3 reflexive) and nequilibrium */
void   orc_ref1d_update(const fvwb1d_params* p, const double* u, const double* w_eq, double* dudt);
/* :553-747 compute_update_sr, the 'WB1' hydrostatic reconstruction */
void   orc_ref1d_update_sr(const fvwb1d_params* p, const double* u, double* dudt);
/* :454-549 compute_update_fvm, the plain 'FVM' solver */
void   orc_ref1d_update_fvm(const fvwb1d_params* p, const double* u, double* dudt);
/* compute_max_speed (:157-169) */
double orc_ref1d_max_speed(const fvwb1d_params* p, const double* u);
/* fvm.f90:188-251 compute_update of the stand-alone 1D program (bc 1 or 2 from
 * p, source 1 none or 2 gravity) */
void   orc_fvm_update(const fvwb1d_params* p, int source, const double* u, double* dudt);

/* ---- 1D configured scheme (deviation form, plain FP64) -------------------- */
void   orc_fv1d_equilibrium(const fvwb1d_params* p, double x, double w[3]);
void   orc_fv1d_face_flux(const fvwb1d_params* p, const double ueq[3], const double dl[3],
                          const double dr[3], double out[3]);
void   orc_fv1d_residual(const fvwb1d_params* p, const double* u, const double* w_eq, double* dudt);
/* dt = cfl*dx/cmax, last step clipped to t_end; runs while t < t_end and
 * steps < max_steps */
void   orc_fv1d_advance(const fvwb1d_params* p, double* u, const double* w_eq, double t_end,
                        int max_steps, double* t_out, int* steps_out, double* dt_last);

/* ---- 2D modal DG ------------------------------------------------------------ */
/* 2d/legendre.f90: legendre (:1-25), legendre_prime (:27-50), gl_quadrature (:77-108) */
double orc_dg_legendre(double x, int n);
double orc_dg_legendre_prime(double x, int n);
void   orc_dg_gl_quadrature(int n, double* xq, double* wq);
/* 2d/benchmark_2d_dg.f90:497-592 tensor-basis projections, AoS(4,nx,ny,order,order) */
void   orc_ref_dg2d_modes_from_nodes(int nx, int ny, int order, const double* nodes, double* modes);
void   orc_ref_dg2d_nodes_from_modes(int nx, int ny, int order, const double* modes, double* nodes);
/* 2d/limiters.f90:312-654: the reference's positivity limiter (limiter_type
 * 'ONP', the shipped one) as written, on its tensor modes AoS(4,nx,ny,o,o) */
void   orc_ref_dg2d_gll_points(int n, double* x);
double orc_ref_dg2d_solve_for_t(double gamma, double eps, const double u[4], const double u_avg[4]);
void   orc_ref_dg2d_positivity(int nx, int ny, int order, double gamma, double eps, double* u);
/* right (columns) and left eigenvectors of the x (dir 0) or y (dir 1) Euler
 * flux Jacobian at the conservative state u, row-major 4x4, L = R^-1 */
void   orc_euler_eigenvectors(double gamma, int dir, const double u[4], double R[16], double L[16]);
/* configured scheme; modes AoS(4,nx,ny,nmodes) of the deviation */
int    orc_dg2d_nmodes(int degree);
void   orc_dg2d_residual(const fvwb_dg_params* p, const double* d, double* L);
void   orc_dg2d_limit(const fvwb_dg_params* p, double* d);
double orc_dg2d_max_rate(const fvwb_dg_params* p, const double* d);
void   orc_dg2d_ssp54_coeffs(double c[16]);
void   orc_dg2d_advance(const fvwb_dg_params* p, double* d, double t_end, int max_steps, double* t_out,
                        int* steps_out, double* dt_last);
/* deviation at the (k+1)^2 volume points, AoS(4,nx,ny,(k+1)^2), point a + (k+1)*b */
void   orc_dg2d_nodal(const fvwb_dg_params* p, const double* d, double* nodal);
void   orc_dg2d_project(const fvwb_dg_params* p, const double* nodal, double* d);

/* ---- 1D modal DG (dg_with_source.f90 with legendre.f90) ------------------- */
double orc_dg1d_legendre(double x, int n);                  /* legendre.f90:1-24   */
double orc_dg1d_legendre_prime(double x, int n);            /* legendre.f90:26-48  */
void   orc_dg1d_gl_quadrature(int n, double* x, double* w); /* legendre.f90:77-127 */
/* the tables of one set-up: Gauss points and weights, phi_i(xi_j) and
 * phi_i'(xi_j) at [i*n+j], phi_i(-1), phi_i(+1) at [2i], [2i+1] */
void   orc_dg1d_tables(int n, double* chsi, double* wq, double* phi_q, double* dphi_q, double* phi_lr);
/* compute_update_exact_delta (:1749-2031); modes and u_eq AoS(3,n,nx) */
void   orc_dg1d_update_exact_delta(const fvwb_dg1d_params* p, const double* delta_u, const double* u_eq,
                                   double* dudt);
/* the 'RKi' SSP(5,4) coefficients as FP32 literals (:264-288) */
void   orc_dg1d_rki_coeffs(double c[15]);
/* the main program's set-up for ninit 7 or 8 (:27-31, 105-125, 148-167) */
void   orc_dg1d_initial(const fvwb_dg1d_params* p, int ninit, double pert, double* delta_u, double* u_eq,
                        double* u_init);
/* the main loop with integrator 'RKi' (:191-199, 264-288, 303-334); returns the
 * number of steps, updates delta_u and u_init (the nodal reconstruction) */
int    orc_dg1d_run_rki(const fvwb_dg1d_params* p, double* delta_u, const double* u_eq, double* u_init,
                        double t_end, int max_steps, double* t_out, double* ts, double* dts);

/* ---- 2D residual distribution (benchmark_rd.f90) --------------------------- */
/* the values the reference reads from its (unshipped) module parameters_rd */
typedef struct {
  int nx, ny;          /* nodes per side (x_i = (i-1) boxlen_x/(nx-1)) */
  int bc;              /* 2 "periodic" copies, 3 "transmissive" copies, 0 none (oracle only) */
  int pad;
  double gamma, boxlen_x, boxlen_y, cfl;
} orc_rd_params;
void   orc_rd_primitive(double gamma, const double u[4], double w[4]);          /* :521-532 */
void   orc_rd_conservative(double gamma, const double w[4], double u[4]);       /* :535-547 */
void   orc_rd_flux(double gamma, const double u[4], double f1[4], double f2[4]); /* :549-571 */
void   orc_rd_source(double gamma, const double ww[4], double s[4]);            /* :747-767 */
double orc_rd_speed(double gamma, const double u[4]);                           /* :478-493 */
double orc_rd_max_speed(const orc_rd_params* p, const double* u);               /* :448-475 */
/* corner residuals 11, 12, 21, 22 of cell (i, j) (:667-708) */
void   orc_rd_cell(const orc_rd_params* p, const double* u, const double* diff, double dt, int i, int j,
                   double phi[4][4]);
/* compute_update_rd (:612-745); u, diff, dudt AoS(4,nx,ny); -1 on bad parameters */
int    orc_rd_update(const orc_rd_params* p, const double* u, const double* diff, double dt, double* dudt);
/* one 'RD1' step of evolve (:346-364); returns dt, -1 on bad parameters */
double orc_rd_step(const orc_rd_params* p, double* u, double t, double t_end);
int    orc_rd_run(const orc_rd_params* p, double* u, double t_end, int max_steps, double* t_out, int* n_out);
/* get_initial_conditions cases 1, 2, 8 (:53-265); -1 for another case */
int    orc_rd_initial(const orc_rd_params* p, int ninit, double eta, double* u);

/* ---- the reference's own 2D DG program (2d/benchmark_2d_dg.f90), as shipped ---- */
int    orc_refdg_valid(const fvwb_dgq_params* p);
/* compute_update (:1137-1479): modes, x, y -> dudt */
void   orc_refdg_update(const fvwb_dgq_params* p, const double* delta_u, const double* x, const double* y,
                        double* dudt);
/* compute_max_speed (:826-870) on u(1:4,1:nx,1:ny): out = (cs_max, v_xmax, v_ymax, speed_max) */
void   orc_refdg_max_speed(const fvwb_dgq_params* p, const double* u, double out[4]);
/* evolve's time step (:668-671) from the means of the modes; ms = compute_max_speed's outputs */
double orc_refdg_dt(const fvwb_dgq_params* p, const double* modes, double t, double t_end, double ms[4]);
/* apply_limiter (:1516-1555), 'ONP' */
void   orc_refdg_limiter(const fvwb_dgq_params* p, double* u);
/* evolve's loop (:666-769) from a limited state; returns the steps taken */
int    orc_refdg_evolve(const fvwb_dgq_params* p, double* delta_u, const double* x, const double* y,
                        double t_end, int max_steps, double* t_io, double* dts);
/* compute_error (:23-89): out = (linf[4], l1[4], l2[4]) */
void   orc_refdg_error(int nx, int ny, int order, const double* u, const double* u_init, double out[12]);

#ifdef __cplusplus
}
#endif
#endif
