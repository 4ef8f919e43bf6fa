/*
 * fvwb_oracle_fv1d.c — CPU FP64 oracle, 1D finite volume (TEST INFRASTRUCTURE).
 *
 * orc_ref1d_* restate benchmark_1d.f90 statement by statement (operation order
 * of the Fortran expressions, no FMA: built with -ffp-contract=off), pinned
 * against the flang build of the reference.  orc_fv1d_* are the configured
 * scheme (deviation form, MUSCL-minmod, Rusanov, SSP-RK2/RK3, walls), pinned by
 * invariants in tests/.  Not linked by the product.
 */
#include "fvwb_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define A3(a, v, i) ((a)[(size_t)(v) + 3u * (size_t)(i)])

/* benchmark_1d.f90:115-123 compute_primitive */
void orc_ref1d_primitive(double gamma, const double u[3], double w[3])
{
  w[0] = u[0];
  w[1] = u[1] / w[0];
  w[2] = (gamma - 1.0) * (u[2] - 0.5 * w[0] * (w[1] * w[1]));
}

/* benchmark_1d.f90:125-132 compute_conservative */
void orc_ref1d_conservative(double gamma, const double w[3], double u[3])
{
  u[0] = w[0];
  u[1] = w[0] * w[1];
  u[2] = w[2] / (gamma - 1.0) + 0.5 * w[0] * (w[1] * w[1]);
}

/* benchmark_1d.f90:171-184 compute_speed: |v| + c */
double orc_ref1d_speed(double gamma, const double u[3])
{
  double w[3];
  orc_ref1d_primitive(gamma, u, w);
  double cs = sqrt(gamma * fmax(w[2], 1e-10) / fmax(w[0], 1e-10));
  return fabs(w[1]) + cs;
}

/* benchmark_1d.f90:186-198 compute_flux */
void orc_ref1d_flux(double gamma, const double u[3], double f[3])
{
  double w[3];
  orc_ref1d_primitive(gamma, u, w);
  f[0] = w[1] * u[0];
  f[1] = w[1] * u[1] + w[2];
  f[2] = w[1] * u[2] + w[2] * w[1];
}

/* benchmark_1d.f90:438-451 compute_llflux(uleft, uright, f_left, f_right) */
void orc_ref1d_llflux(double gamma, const double ul[3], const double ur[3], const double fl[3],
                      const double fr[3], double out[3])
{
  double cl = orc_ref1d_speed(gamma, ul), cr = orc_ref1d_speed(gamma, ur);
  double cmax = fmax(cl, cr);
  for (int k = 0; k < 3; ++k) out[k] = 0.5 * (fr[k] + fl[k]) - 0.5 * cmax * (ur[k] - ul[k]);
}

/* benchmark_1d.f90:134-155 get_equilibrium_solution at one point, written for
 * a general g (the reference has g = 1: the factor `*1*` of :151-153). */
static void equilibrium_g(double gamma, int neq, double g, double x, double w[3])
{
  if (neq == 3) {
    double base = 1 - ((gamma - 1) / gamma) * g * x;
    w[0] = pow(base, 1 / (gamma - 1));
    w[1] = 0;
    w[2] = pow(base, gamma / (gamma - 1));
  } else {
    w[0] = exp(-(g * x));
    w[1] = 0;
    w[2] = exp(-(g * x));
  }
}

void orc_ref1d_equilibrium(double gamma, int neq, int n, const double* x, double* w)
{
  for (int i = 0; i < n; ++i) {
    if (neq == 3) {
      double base = 1 - ((gamma - 1) / gamma) * 1 * x[i];
      A3(w, 0, i) = pow(base, 1 / (gamma - 1));
      A3(w, 1, i) = 0;
      A3(w, 2, i) = pow(base, gamma / (gamma - 1));
    } else {
      A3(w, 0, i) = exp(-x[i]);
      A3(w, 1, i) = 0;
      A3(w, 2, i) = exp(-x[i]);
    }
  }
}

/* benchmark_1d.f90:263-377 compute_update (solver 'EQL'): exact deviation
 * residual, first order, Rusanov; see the comments on the boundary faces. */
void orc_ref1d_update(const fvwb1d_params* p, const double* u, const double* w_eq, double* dudt)
{
  const int nx = p->nx, nf = nx + 1;
  const double gamma = p->gamma;
  const double dx = p->boxlen / (double)nx;
  const double oneoverdx = 1 / dx;
  double* delta = malloc(sizeof(double) * 3 * nx);
  double* uleft = malloc(sizeof(double) * 3 * nx);
  double* uright = malloc(sizeof(double) * 3 * nx);
  double* fleft = malloc(sizeof(double) * 3 * nx);
  double* fright = malloc(sizeof(double) * 3 * nx);
  double* ueqc = malloc(sizeof(double) * 3 * nx);
  double* xf = malloc(sizeof(double) * nf);
  double* xc = malloc(sizeof(double) * nx);
  double* weqf = malloc(sizeof(double) * 3 * nf);
  double* ueqf = malloc(sizeof(double) * 3 * nf);
  double* feq = malloc(sizeof(double) * 3 * nf);
  double* fr = malloc(sizeof(double) * 3 * nf);

  for (int i = 0; i < nx; ++i) orc_ref1d_conservative(gamma, &w_eq[3 * i], &ueqc[3 * i]);
  for (int k = 0; k < 3 * nx; ++k) delta[k] = u[k] - ueqc[k];
  for (int i = 0; i < nx; ++i) xc[i] = ((double)(i + 1) - 0.5) * dx;   /* (i-0.5)*dx, exact in FP32 */
  for (int i = 0; i < nf; ++i) xf[i] = (double)i * dx;                 /* (i-1)*dx */
  orc_ref1d_equilibrium(gamma, p->nequilibrium, nf, xf, weqf);
  for (int i = 0; i < nf; ++i) orc_ref1d_conservative(gamma, &weqf[3 * i], &ueqf[3 * i]);
  for (int i = 0; i < nx; ++i)
    for (int k = 0; k < 3; ++k) {
      A3(uleft, k, i) = A3(delta, k, i) + A3(ueqf, k, i);
      A3(uright, k, i) = A3(delta, k, i) + A3(ueqf, k, i + 1);
    }
  for (int i = 0; i < nx; ++i) {
    orc_ref1d_flux(gamma, &uleft[3 * i], &fleft[3 * i]);
    orc_ref1d_flux(gamma, &uright[3 * i], &fright[3 * i]);
  }
  for (int i = 0; i < nf; ++i) orc_ref1d_flux(gamma, &ueqf[3 * i], &feq[3 * i]);

  /* faces (:311-368): zero-based face f sits between cells f-1 and f */
  for (int f = 0; f < nf; ++f) {
    int il = f - 1, ir = f;
    if (p->bc == FVWB1D_BC_PERIODIC) {
      if (f == 0) il = nx - 1;
      if (f == nx) ir = 0;
    } else {
      if (f == 0) il = 0;
      if (f == nx) ir = nx - 1;
    }
    orc_ref1d_llflux(gamma, &uright[3 * il], &uleft[3 * ir], &fright[3 * il], &fleft[3 * ir], &fr[3 * f]);
    if (p->bc == FVWB1D_BC_REFLEXIVE && f == 0) {
      /* :341-351: equilibrium (1,0,1) on the outside, plus delta of cell 1 */
      const double wm[3] = {1., 0., 1.};
      double uf[3], s[3], fm[3];
      orc_ref1d_conservative(gamma, wm, uf);
      for (int k = 0; k < 3; ++k) s[k] = uf[k] + delta[k];
      orc_ref1d_flux(gamma, s, fm);
      orc_ref1d_llflux(gamma, s, &uleft[3 * ir], fm, &fleft[3 * ir], &fr[3 * f]);
    }
    if (p->bc == FVWB1D_BC_REFLEXIVE && f == nx) {
      /* :352-361 */
      const double wp[3] = {1., 0., 1.};
      double up[3], s[3], l[3], fp[3];
      orc_ref1d_conservative(gamma, wp, up);
      for (int k = 0; k < 3; ++k) s[k] = up[k] + A3(delta, k, nx - 1);
      orc_ref1d_flux(gamma, s, fp);
      for (int k = 0; k < 3; ++k) l[k] = A3(uright, k, il) + A3(delta, k, nx - 1);
      orc_ref1d_llflux(gamma, l, up, &fright[3 * il], fp, &fr[3 * f]);
    }
  }

  /* get_source (:379-405) is called with the conservative states u and u_eq */
  const double delta_s = 1 / (double)nx;
  for (int i = 0; i < nx; ++i) {
    double xm = i == 0 ? xc[0] - delta_s : xc[i - 1];
    double xp = i == nx - 1 ? xc[nx - 1] + delta_s : xc[i + 1];
    double s[3], se[3];
    s[0] = 0;
    s[1] = -(A3(u, 0, i) * 1 * (xp - xm) / (2 * delta_s));
    s[2] = -(A3(u, 0, i) * A3(u, 1, i) * 1 * (xp - xm) / (2 * delta_s));
    se[0] = 0;
    se[1] = -(A3(ueqc, 0, i) * 1 * (xp - xm) / (2 * delta_s));
    se[2] = -(A3(ueqc, 0, i) * A3(ueqc, 1, i) * 1 * (xp - xm) / (2 * delta_s));
    for (int k = 0; k < 3; ++k)
      A3(dudt, k, i) = -((A3(fr, k, i + 1) - A3(fr, k, i)) * oneoverdx) + s[k] +
                       (A3(feq, k, i + 1) - A3(feq, k, i)) * oneoverdx - se[k];
  }
  for (int k = 0; k < 3; ++k) {
    A3(dudt, k, 0) = A3(dudt, k, 1);
    A3(dudt, k, nx - 1) = A3(dudt, k, nx - 2);
  }
  free(delta); free(uleft); free(uright); free(fleft); free(fright); free(ueqc);
  free(xf); free(xc); free(weqf); free(ueqf); free(feq); free(fr);
}

/* compute_update_sr (benchmark_1d.f90:553-747), the 'WB1' solver the reference
 * ships (parameters.f90:8): hydrostatic reconstruction of each cell's face states
 * from its entropy K = p/rho^gamma and enthalpy h = gamma/(gamma-1) p/rho, moved
 * along phi = x (:596-629); Rusanov between the reconstructed states at the
 * interior faces (:637-665; the loop runs over faces 2..nx only, so the boundary
 * branches inside it never execute); the source of get_source_rg (:409-434); the
 * boundary cells copy their neighbours (:741-742).  Constants as written: the
 * 1e-5 floors are default-real literals, oneoverdx = boxlen/dx (:573). */
void orc_ref1d_update_sr(const fvwb1d_params* p, const double* u, double* dudt)
{
  const int nx = p->nx;
  const double gamma = p->gamma;
  const double dx = p->boxlen / (double)nx;
  const double oneoverdx = p->boxlen / dx;
  const double e5 = (double)1e-5f;
  double* wl = malloc(sizeof(double) * 3 * nx);
  double* wr = malloc(sizeof(double) * 3 * nx);
  double* ul = malloc(sizeof(double) * 3 * nx);
  double* ur = malloc(sizeof(double) * 3 * nx);
  double* fl = malloc(sizeof(double) * 3 * nx);
  double* fr = malloc(sizeof(double) * 3 * nx);
  double* w = malloc(sizeof(double) * 3 * nx);
  double* x = malloc(sizeof(double) * nx);
  double* F = malloc(sizeof(double) * 3 * (nx + 1));
  for (int i = 0; i < nx; ++i) {
    x[i] = ((double)(i + 1) - 0.5) * dx;                      /* (i-0.5)*dx, exact in FP32 */
    orc_ref1d_primitive(gamma, &u[3 * i], &w[3 * i]);
    const double rho = w[3 * i], pr = w[3 * i + 2];
    const double h = fmax(pr, e5) / fmax(e5, rho) * (1 + 1.0 / (gamma - 1));
    const double h0l = h + x[i] - (double)i * dx;             /* phi = x at the centre and faces */
    const double h0r = h + x[i] - (double)(i + 1) * dx;
    const double kap = fmax(e5, pr) / pow(fmax(e5, rho), gamma);
    A3(wl, 0, i) = pow((1.0 / kap) * (gamma - 1) / gamma * h0l, 1 / (gamma - 1));
    A3(wl, 1, i) = w[3 * i + 1];
    A3(wl, 2, i) = pow(1.0 / kap, 1 / (gamma - 1)) * pow((gamma - 1) / gamma * h0l, gamma / (gamma - 1));
    A3(wr, 0, i) = pow((1.0 / kap) * (gamma - 1) / gamma * h0r, 1 / (gamma - 1));
    A3(wr, 1, i) = w[3 * i + 1];
    A3(wr, 2, i) = pow(1.0 / kap, 1 / (gamma - 1)) * pow((gamma - 1) / gamma * h0r, gamma / (gamma - 1));
    orc_ref1d_conservative(gamma, &wl[3 * i], &ul[3 * i]);
    orc_ref1d_conservative(gamma, &wr[3 * i], &ur[3 * i]);
    orc_ref1d_flux(gamma, &ul[3 * i], &fl[3 * i]);
    orc_ref1d_flux(gamma, &ur[3 * i], &fr[3 * i]);
  }
  /* interior faces: zero-based face f between cells f-1 and f */
  for (int f = 1; f < nx; ++f)
    orc_ref1d_llflux(gamma, &ur[3 * (f - 1)], &ul[3 * f], &fr[3 * (f - 1)], &fl[3 * f], &F[3 * f]);
  /* get_source_rg with delta = 1/nx and x extended by delta at both ends */
  const double delta = 1.0 / (double)nx;
  for (int i = 1; i < nx - 1; ++i) {
    const double xm = i == 0 ? x[0] - delta : x[i - 1];
    const double xp = i == nx - 1 ? x[nx - 1] + delta : x[i + 1];
    const double s1 = (A3(wr, 2, i) - A3(wl, 2, i)) / delta;
    const double s2 = -(w[3 * i] * w[3 * i + 1] * 1 * (xp - xm) / (2 * delta));
    dudt[3 * i + 0] = -(F[3 * (i + 1) + 0] - F[3 * i + 0]) * oneoverdx + 0.0;
    dudt[3 * i + 1] = -(F[3 * (i + 1) + 1] - F[3 * i + 1]) * oneoverdx + s1;
    dudt[3 * i + 2] = -(F[3 * (i + 1) + 2] - F[3 * i + 2]) * oneoverdx + s2;
  }
  for (int k = 0; k < 3; ++k) {
    dudt[k] = dudt[3 + k];
    dudt[3 * (nx - 1) + k] = dudt[3 * (nx - 2) + k];
  }
  free(wl); free(wr); free(ul); free(ur); free(fl); free(fr); free(w); free(x); free(F);
}

/* compute_update_fvm (benchmark_1d.f90:454-549), the plain 'FVM' solver: Rusanov
 * between the cell states themselves, get_source (:379-405) on the conservative
 * state, and the boundary cells copy their neighbours (:546-547), so only the
 * interior faces reach dudt and the bc code does not.  oneoverdx = 1/dx (:470). */
void orc_ref1d_update_fvm(const fvwb1d_params* p, const double* u, double* dudt)
{
  const int nx = p->nx;
  const double gamma = p->gamma;
  const double dx = p->boxlen / (double)nx;
  const double oneoverdx = 1 / dx;
  const double delta = 1 / (double)nx;
  double* f = malloc(sizeof(double) * 3 * nx);
  double* F = malloc(sizeof(double) * 3 * (nx + 1));
  for (int i = 0; i < nx; ++i) orc_ref1d_flux(gamma, &u[3 * i], &f[3 * i]);
  for (int k = 1; k < nx; ++k) orc_ref1d_llflux(gamma, &u[3 * (k - 1)], &u[3 * k], &f[3 * (k - 1)], &f[3 * k], &F[3 * k]);
  for (int i = 1; i < nx - 1; ++i) {
    const double xm = ((double)i - 0.5) * dx;        /* x(i-1) of the one-based cell i+1 */
    const double xp = ((double)(i + 2) - 0.5) * dx;
    const double s[3] = {0.0, -(A3(u, 0, i) * 1 * (xp - xm) / (2 * delta)),
                         -(A3(u, 0, i) * A3(u, 1, i) * 1 * (xp - xm) / (2 * delta))};
    for (int k = 0; k < 3; ++k) A3(dudt, k, i) = -((A3(F, k, i + 1) - A3(F, k, i)) * oneoverdx) + s[k];
  }
  for (int k = 0; k < 3; ++k) {
    A3(dudt, k, 0) = A3(dudt, k, 1);
    A3(dudt, k, nx - 1) = A3(dudt, k, nx - 2);
  }
  free(f); free(F);
}

/* compute_update of the stand-alone program fvm.f90 (:188-251), with its
 * compute_source (:253-264) and compute_llflux (:271-288): Rusanov between cell
 * states, each cell computing both of its faces; bc 1 wraps the neighbours of
 * the end cells, bc 2 repeats the end cell (:200-230); source 2 is
 * -rho*(x+ - x-)/(2 dx) on momentum and -rho*v*(x+ - x-)/(2 dx) on energy from
 * the primitive state, with x- and x+ clamped to the cell's own centre at the two
 * ends (:232-243), source 1 is zero.  No boundary copy: every cell is its own.
 * oneoverdx = 1.0/dx (:198). */
void orc_fvm_update(const fvwb1d_params* p, int source, const double* u, double* dudt)
{
  const int nx = p->nx;
  const double gamma = p->gamma;
  const double dx = p->boxlen / (double)nx;
  const double oneoverdx = 1.0 / dx;
  for (int i = 0; i < nx; ++i) {
    int im = i - 1, ip = i + 1;
    if (p->bc == FVWB1D_BC_PERIODIC) {
      if (i == 0) im = nx - 1;
      if (i == nx - 1) ip = 0;
    } else {
      if (i == 0) im = 0;
      if (i == nx - 1) ip = nx - 1;
    }
    double fm[3], fc[3], fp[3], FL[3], FR[3];
    orc_ref1d_flux(gamma, &u[3 * im], fm);
    orc_ref1d_flux(gamma, &u[3 * i], fc);
    orc_ref1d_flux(gamma, &u[3 * ip], fp);
    orc_ref1d_llflux(gamma, &u[3 * im], &u[3 * i], fm, fc, FL);
    orc_ref1d_llflux(gamma, &u[3 * i], &u[3 * ip], fc, fp, FR);
    double s[3] = {0.0, 0.0, 0.0};
    if (source == 2) {
      double xm = ((double)(i + 1 - 1) - 0.5) * dx, xp = ((double)(i + 1 + 1) - 0.5) * dx;
      if (i == 0) xm = ((double)1 - 0.5) * dx;
      if (i == nx - 1) xp = ((double)nx - 0.5) * dx;
      double w[3];
      orc_ref1d_primitive(gamma, &u[3 * i], w);
      s[1] = -(w[0] * 1 * (xp - xm) / (2 * dx));
      s[2] = -(w[0] * w[1] * 1 * (xp - xm) / (2 * dx));
    }
    for (int k = 0; k < 3; ++k) A3(dudt, k, i) = -(oneoverdx * (FR[k] - FL[k])) + s[k];
  }
}

/* benchmark_1d.f90:157-169 compute_max_speed */
double orc_ref1d_max_speed(const fvwb1d_params* p, const double* u)
{
  double cmax = 0.0;
  for (int i = 0; i < p->nx; ++i) cmax = fmax(cmax, orc_ref1d_speed(p->gamma, &u[3 * i]));
  return cmax;
}

/* ---- configured scheme --------------------------------------------------- */

void orc_fv1d_equilibrium(const fvwb1d_params* p, double x, double w[3])
{
  equilibrium_g(p->gamma, p->nequilibrium, p->g, x, w);
}

/* deviation of cell i (any integer) with the ghost rule of p->bc */
static void dev1d(const fvwb1d_params* p, const double* d, int i, double out[3])
{
  const int nx = p->nx;
  int src = i, mirror = 0;
  if (i < 0) {
    if (p->bc == FVWB1D_BC_WALL) { src = -1 - i; mirror = 1; } else src = 0;
  } else if (i >= nx) {
    if (p->bc == FVWB1D_BC_WALL) { src = 2 * nx - 1 - i; mirror = 1; } else src = nx - 1;
  }
  out[0] = A3(d, 0, src);
  out[1] = mirror ? -A3(d, 1, src) : A3(d, 1, src);
  out[2] = A3(d, 2, src);
}

/* numerical flux minus equilibrium flux at a face with equilibrium ueq and
 * reconstructed deviations dl (left) and dr (right) */
void orc_fv1d_face_flux(const fvwb1d_params* p, const double ueq[3], const double dl[3], const double dr[3],
                        double out[3])
{
  double ul[3], ur[3], fl[3], fr[3], fe[3], f[3];
  for (int k = 0; k < 3; ++k) { ul[k] = ueq[k] + dl[k]; ur[k] = ueq[k] + dr[k]; }
  orc_ref1d_flux(p->gamma, ul, fl);
  orc_ref1d_flux(p->gamma, ur, fr);
  orc_ref1d_flux(p->gamma, ueq, fe);
  orc_ref1d_llflux(p->gamma, ul, ur, fl, fr, f);
  for (int k = 0; k < 3; ++k) out[k] = f[k] - fe[k];
}

/* residual on deviations d (AoS(3,nx)); ueqf: conservative face equilibrium (nx+1) */
static void residual_dev(const fvwb1d_params* p, const double* d, const double* ueqf, double* L)
{
  const int nx = p->nx;
  const double dx = p->boxlen / (double)nx, odx = 1.0 / dx;
  double* G = malloc(sizeof(double) * 3 * (nx + 1));
  for (int f = 0; f <= nx; ++f) {
    double a[3], b[3], dl[3], dr[3];
    dev1d(p, d, f - 1, a);
    dev1d(p, d, f, b);
    if (p->recon == FVWB_RECON_MUSCL_MINMOD) {
      double am[3], bp[3];
      dev1d(p, d, f - 2, am);
      dev1d(p, d, f + 1, bp);
      for (int k = 0; k < 3; ++k) {
        double sa = orc_minmod(a[k] - am[k], b[k] - a[k], b[k] - a[k]);
        double sb = orc_minmod(b[k] - a[k], bp[k] - b[k], bp[k] - b[k]);
        dl[k] = a[k] + 0.5 * sa;
        dr[k] = b[k] - 0.5 * sb;
      }
    } else {
      memcpy(dl, a, sizeof a);
      memcpy(dr, b, sizeof b);
    }
    orc_fv1d_face_flux(p, &ueqf[3 * f], dl, dr, &G[3 * f]);
  }
  for (int i = 0; i < nx; ++i) {
    const double s[3] = {0.0, -p->g * A3(d, 0, i), -p->g * A3(d, 1, i)};
    for (int k = 0; k < 3; ++k) A3(L, k, i) = -(A3(G, k, i + 1) - A3(G, k, i)) * odx + s[k];
  }
  free(G);
}

static double* face_table(const fvwb1d_params* p)
{
  const int nx = p->nx;
  const double dx = p->boxlen / (double)nx;
  double* t = malloc(sizeof(double) * 3 * (nx + 1));
  for (int f = 0; f <= nx; ++f) {
    double w[3];
    orc_fv1d_equilibrium(p, (double)f * dx, w);
    orc_ref1d_conservative(p->gamma, w, &t[3 * f]);
  }
  return t;
}

void orc_fv1d_residual(const fvwb1d_params* p, const double* u, const double* w_eq, double* dudt)
{
  const int nx = p->nx;
  double* d = malloc(sizeof(double) * 3 * nx);
  for (int i = 0; i < nx; ++i) {
    double ue[3];
    orc_ref1d_conservative(p->gamma, &w_eq[3 * i], ue);
    for (int k = 0; k < 3; ++k) A3(d, k, i) = A3(u, k, i) - ue[k];
  }
  double* t = face_table(p);
  residual_dev(p, d, t, dudt);
  free(t);
  free(d);
}

/* max over cells of |v| + c of U_eq(centre) + d */
static double max_speed_dev(const fvwb1d_params* p, const double* ueqc, const double* d)
{
  double cmax = 0.0;
  for (int i = 0; i < p->nx; ++i) {
    double s[3];
    for (int k = 0; k < 3; ++k) s[k] = A3(ueqc, k, i) + A3(d, k, i);
    cmax = fmax(cmax, orc_ref1d_speed(p->gamma, s));
  }
  return cmax;
}

/* Steps the deviation d = u - U_eq(w_eq): dt = cfl*dx/cmax, the last step
 * clipped to t_end; SSP-RK2 as benchmark_1d.f90:230-235 combines its stages
 * (u = 0.5*u + 0.5*w1 + 0.5*dt*L(w1)), SSP-RK3 in Shu-Osher form. */
void orc_fv1d_advance(const fvwb1d_params* p, double* u, const double* w_eq, double t_end, int max_steps,
                      double* t_out, int* steps_out, double* dt_last)
{
  const int nx = p->nx, n3 = 3 * nx;
  const double dx = p->boxlen / (double)nx;
  double* ueqc = malloc(sizeof(double) * n3);
  double* d = malloc(sizeof(double) * n3);
  double* d1 = malloc(sizeof(double) * n3);
  double* d2 = malloc(sizeof(double) * n3);
  double* L = malloc(sizeof(double) * n3);
  double* tab = face_table(p);
  for (int i = 0; i < nx; ++i) orc_ref1d_conservative(p->gamma, &w_eq[3 * i], &ueqc[3 * i]);
  for (int k = 0; k < n3; ++k) d[k] = u[k] - ueqc[k];
  double t = 0.0, dt = 0.0;
  int iter = 0;
  while (t < t_end && iter < max_steps) {
    double cmax = max_speed_dev(p, ueqc, d);
    dt = p->cfl * dx / cmax;
    if (t + dt > t_end) dt = t_end - t;
    residual_dev(p, d, tab, L);
    for (int k = 0; k < n3; ++k) d1[k] = d[k] + dt * L[k];
    residual_dev(p, d1, tab, L);
    if (p->integrator == FVWB_SSPRK3) {
      for (int k = 0; k < n3; ++k) d2[k] = 0.75 * d[k] + 0.25 * d1[k] + 0.25 * dt * L[k];
      residual_dev(p, d2, tab, L);
      for (int k = 0; k < n3; ++k) d[k] = (1.0 / 3.0) * d[k] + (2.0 / 3.0) * d2[k] + (2.0 / 3.0) * dt * L[k];
    } else {
      for (int k = 0; k < n3; ++k) d[k] = 0.5 * d[k] + 0.5 * d1[k] + 0.5 * dt * L[k];
    }
    t = t + dt;
    iter = iter + 1;
  }
  for (int k = 0; k < n3; ++k) u[k] = ueqc[k] + d[k];
  if (t_out) *t_out = t;
  if (steps_out) *steps_out = iter;
  if (dt_last) *dt_last = dt;
  free(ueqc); free(d); free(d1); free(d2); free(L); free(tab);
}
