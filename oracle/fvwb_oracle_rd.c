/*
 * fvwb_oracle_rd.c — CPU FP64 oracle, 2D residual distribution (TEST
 * INFRASTRUCTURE).
 *
 * Restates benchmark_rd.f90 statement by statement: the Fortran expression
 * order, no FMA (built with -ffp-contract=off), default-real literals (0.5,
 * 0.25, 1.21, 0.3, 100*1.21) rounded as the reference compiles them.  The
 * reference program needs a module parameters_rd that it does not ship, so it
 * cannot be built here; these functions are pinned by the closed forms and
 * invariants in tests/test_oracle_rd.py (uniform states, per-cell conservation
 * of the distributed residual, the corner differences that isolate the
 * Lax-Friedrichs term, transposition, linearity in diff).  Not linked by the
 * product.
 *
 * Layout: u(1:4, 1:nx, 1:ny) of the reference with mx = my = 1, i.e.
 * u[4*(j*nx + i) + v].
 */
#include "fvwb_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define U4(a, v, i, j, nx) ((a)[4u * ((size_t)(j) * (size_t)(nx) + (size_t)(i)) + (size_t)(v)])

/* benchmark_rd.f90:521-532 compute_primitive */
void orc_rd_primitive(double gamma, const double u[4], double w[4])
{
  w[0] = u[0];
  w[1] = u[1] / u[0];
  w[2] = u[2] / u[0];
  w[3] = (gamma - 1.0) * (u[3] - 0.5 * w[0] * (w[1] * w[1] + w[2] * w[2]));
}

/* benchmark_rd.f90:535-547 compute_conservative */
void orc_rd_conservative(double gamma, const double w[4], double u[4])
{
  u[0] = w[0];
  u[1] = w[0] * w[1];
  u[2] = w[0] * w[2];
  u[3] = w[3] / (gamma - 1.0) + 0.5 * (w[0] * (w[1] * w[1] + w[2] * w[2]));
}

/* benchmark_rd.f90:549-571 compute_flux: f1 along x, f2 along y */
void orc_rd_flux(double gamma, const double u[4], double f1[4], double f2[4])
{
  double w[4];
  orc_rd_primitive(gamma, u, w);
  f2[0] = w[0] * w[2];
  f2[1] = w[0] * w[1] * w[2];
  f2[2] = w[2] * u[2] + w[3];
  f2[3] = w[2] * u[3] + w[2] * w[3];
  f1[0] = w[0] * w[1];
  f1[1] = w[1] * u[1] + w[3];
  f1[2] = w[0] * w[1] * w[2];
  f1[3] = w[1] * u[3] + w[1] * w[3];
}

/* benchmark_rd.f90:747-767 get_source with grad_p = (1, 1).  The caller
 * (:681-684) hands it the primitive vector ww, and get_source converts its
 * argument to primitive variables once more (:756), so the energy source is
 * -rho*(vx/rho + vy/rho); the second conversion's pressure is unused. */
void orc_rd_source(double gamma, const double ww[4], double s[4])
{
  double w[4];
  orc_rd_primitive(gamma, ww, w);
  s[0] = 0.0;
  s[1] = -w[0] * 1.0;
  s[2] = -w[0] * 1.0;
  s[3] = -w[0] * (w[1] * 1.0 + w[2] * 1.0);
}

/* benchmark_rd.f90:478-493 compute_speed: |v| + c */
double orc_rd_speed(double gamma, const double u[4])
{
  double w[4];
  orc_rd_primitive(gamma, u, w);
  double cs = sqrt(gamma * fmax(w[3], 1e-10) / fmax(w[0], 1e-10));
  return sqrt(w[1] * w[1] + w[2] * w[2]) + cs;
}

/* benchmark_rd.f90:448-475 compute_max_speed (nodes visited x outer, y inner) */
double orc_rd_max_speed(const orc_rd_params* p, const double* u)
{
  double smax = 0.0;
  for (int i = 0; i < p->nx; ++i)
    for (int j = 0; j < p->ny; ++j) {
      double s = orc_rd_speed(p->gamma, &U4(u, 0, i, j, p->nx));
      if (s >= smax) smax = fmax(smax, s);
    }
  return smax;
}

/* The four corner residuals of cell (i, j) (nodes i..i+1, j..j+1), in the
 * order 11, 12, 21, 22 of benchmark_rd.f90:667-708. */
void orc_rd_cell(const orc_rd_params* p, const double* u, const double* diff, double dt, int i, int j,
                 double phi[4][4])
{
  const int nx = p->nx;
  const double dx = p->boxlen_x / (double)(p->nx - 1);
  const double dy = p->boxlen_y / (double)(p->ny - 1);
  const double* uu11 = &U4(u, 0, i, j, nx);
  const double* uu12 = &U4(u, 0, i + 1, j, nx);
  const double* uu21 = &U4(u, 0, i, j + 1, nx);
  const double* uu22 = &U4(u, 0, i + 1, j + 1, nx);
  double ff11[4], ff12[4], ff21[4], ff22[4], g11[4], g12[4], g21[4], g22[4];
  orc_rd_flux(p->gamma, uu11, ff11, g11);
  orc_rd_flux(p->gamma, uu12, ff12, g12);
  orc_rd_flux(p->gamma, uu21, ff21, g21);
  orc_rd_flux(p->gamma, uu22, ff22, g22);
  double ww11[4], ww12[4], ww21[4], ww22[4];
  orc_rd_primitive(p->gamma, uu11, ww11);
  orc_rd_primitive(p->gamma, uu12, ww12);
  orc_rd_primitive(p->gamma, uu22, ww22);
  orc_rd_primitive(p->gamma, uu21, ww21);
  double s11[4], s12[4], s21[4], s22[4];
  orc_rd_source(p->gamma, ww11, s11);
  orc_rd_source(p->gamma, ww12, s12);
  orc_rd_source(p->gamma, ww21, s21);
  orc_rd_source(p->gamma, ww22, s22);
  double totalphi[4];
  for (int v = 0; v < 4; ++v)
    totalphi[v] = -0.5 * (ff11[v] + ff21[v]) + 0.5 * (ff12[v] + ff22[v]) - 0.5 * (g22[v] + g21[v]) +
                  0.5 * (g11[v] + g12[v]) - 0.25 * (s11[v] + s21[v] + s12[v] + s22[v]) * dx * dy;
  /* compute_speed of the primitive state, as the reference passes ww */
  double sp[4];
  const double* ws[4] = {ww11, ww12, ww21, ww22};
  for (int k = 0; k < 4; ++k) {
    double w2[4];
    orc_rd_primitive(p->gamma, ws[k], w2);
    double cs = sqrt(p->gamma * fmax(w2[3], 1e-10) / fmax(w2[0], 1e-10));
    sp[k] = sqrt(w2[1] * w2[1] + w2[2] * w2[2]) + cs;
  }
  /* maxval([speed1, speed2, speed3, speed4]) */
  double alpha = sp[0];
  for (int k = 1; k < 4; ++k)
    if (sp[k] > alpha) alpha = sp[k];
  const double* d11 = &U4(diff, 0, i, j, nx);
  const double* d12 = &U4(diff, 0, i + 1, j, nx);
  const double* d21 = &U4(diff, 0, i, j + 1, nx);
  const double* d22 = &U4(diff, 0, i + 1, j + 1, nx);
  for (int v = 0; v < 4; ++v) {
    double ubar = 0.25 * (uu11[v] + uu22[v] + uu21[v] + uu12[v]);
    phi[0][v] = dy * dx * d11[v] * 0.25 + dt * (totalphi[v] * 0.25 + alpha * (uu11[v] - ubar) * dy);
    phi[1][v] = dy * dx * d12[v] * 0.25 + dt * (totalphi[v] * 0.25 + alpha * (uu12[v] - ubar) * dy);
    phi[2][v] = dy * dx * d21[v] * 0.25 + dt * (totalphi[v] * 0.25 + alpha * (uu21[v] - ubar) * dx);
    phi[3][v] = dy * dx * d22[v] * 0.25 + dt * (totalphi[v] * 0.25 + alpha * (uu22[v] - ubar) * dx);
  }
}

/* benchmark_rd.f90:612-745 compute_update_rd.  Cells are visited x outer, y
 * inner and scattered into dudt in that order; then the boundary copies of
 * :722-737 (bc 2 "periodic", bc 3 "transmissive", applied in the listed
 * order).  bc 0 skips the copies: a test hook of the oracle only (the
 * reference stops on any bc but 2 and 3). */
int orc_rd_update(const orc_rd_params* p, const double* u, const double* diff, double dt, double* dudt)
{
  const int nx = p->nx, ny = p->ny;
  if (nx < 2 || ny < 2 || (p->bc != 0 && p->bc != 2 && p->bc != 3)) return -1;
  memset(dudt, 0, sizeof(double) * 4u * (size_t)nx * (size_t)ny);
  for (int i = 0; i < nx - 1; ++i)
    for (int j = 0; j < ny - 1; ++j) {
      double phi[4][4];
      orc_rd_cell(p, u, diff, dt, i, j, phi);
      for (int v = 0; v < 4; ++v) {
        U4(dudt, v, i, j, nx) += phi[0][v];
        U4(dudt, v, i + 1, j, nx) += phi[1][v];
        U4(dudt, v, i, j + 1, nx) += phi[2][v];
        U4(dudt, v, i + 1, j + 1, nx) += phi[3][v];
      }
    }
  if (p->bc == 2) {
    for (int j = 0; j < ny; ++j)
      for (int v = 0; v < 4; ++v) U4(dudt, v, 0, j, nx) = U4(dudt, v, nx - 1, j, nx);
    for (int j = 0; j < ny; ++j)
      for (int v = 0; v < 4; ++v) U4(dudt, v, nx - 1, j, nx) = U4(dudt, v, 0, j, nx);
    for (int i = 0; i < nx; ++i)
      for (int v = 0; v < 4; ++v) U4(dudt, v, i, ny - 1, nx) = U4(dudt, v, i, 0, nx);
    for (int i = 0; i < nx; ++i)
      for (int v = 0; v < 4; ++v) U4(dudt, v, i, 0, nx) = U4(dudt, v, i, ny - 1, nx);
  } else if (p->bc == 3) {
    for (int j = 0; j < ny; ++j)
      for (int v = 0; v < 4; ++v) U4(dudt, v, 0, j, nx) = U4(dudt, v, 1, j, nx);
    for (int j = 0; j < ny; ++j)
      for (int v = 0; v < 4; ++v) U4(dudt, v, nx - 1, j, nx) = U4(dudt, v, nx - 2, j, nx);
    for (int i = 0; i < nx; ++i)
      for (int v = 0; v < 4; ++v) U4(dudt, v, i, ny - 1, nx) = U4(dudt, v, i, ny - 2, nx);
    for (int i = 0; i < nx; ++i)
      for (int v = 0; v < 4; ++v) U4(dudt, v, i, 0, nx) = U4(dudt, v, i, 1, nx);
  }
  return 0;
}

/* One step of evolve, solver 'RD1' (benchmark_rd.f90:346-364): dt from the
 * maximum speed, a first pass with diff = 0, then a second pass at the same u
 * with diff = up - u.  Returns the dt taken, or -1 on bad parameters. */
double orc_rd_step(const orc_rd_params* p, double* u, double t, double t_end)
{
  const size_t n = 4u * (size_t)p->nx * (size_t)p->ny;
  const double dx = p->boxlen_x / (double)(p->nx - 1);
  const double dy = p->boxlen_y / (double)(p->ny - 1);
  double cmax = orc_rd_max_speed(p, u);
  double dt = fmin(p->cfl * dy * dx / cmax / (2.0 * 1.0 + 1.0), t_end - t);
  double* diff = calloc(n, sizeof(double));
  double* dudt = malloc(n * sizeof(double));
  double* up = malloc(n * sizeof(double));
  if (!diff || !dudt || !up || orc_rd_update(p, u, diff, dt, dudt) != 0) {
    free(diff); free(dudt); free(up);
    return -1.0;
  }
  for (size_t k = 0; k < n; ++k) up[k] = u[k] - dudt[k] / (dx * dy);
  for (size_t k = 0; k < n; ++k) diff[k] = up[k] - u[k];
  orc_rd_update(p, u, diff, dt, dudt);
  for (size_t k = 0; k < n; ++k) u[k] = up[k] - dudt[k] / (dx * dy);
  free(diff); free(dudt); free(up);
  return dt;
}

/* evolve's loop (benchmark_rd.f90:356-385): steps while t < t_end, at most
 * max_steps of them. */
int orc_rd_run(const orc_rd_params* p, double* u, double t_end, int max_steps, double* t_out, int* n_out)
{
  double t = 0.0;
  int n = 0;
  while (t < t_end && n < max_steps) {
    double dt = orc_rd_step(p, u, t, t_end);
    if (dt < 0.0) return -1;
    t = t + dt;
    ++n;
  }
  *t_out = t;
  *n_out = n;
  return 0;
}

/* benchmark_rd.f90:24-51 get_coords and :53-265 get_initial_conditions, cases
 * 1 (density bump in a uniform flow), 2 (isothermal rest state, rho_0 = 1.21)
 * and 8 (rest state plus a pressure bump of amplitude eta at (0.3, 0.3)),
 * converted with compute_conservative.  Returns -1 for another case. */
int orc_rd_initial(const orc_rd_params* p, int ninit, double eta, double* u)
{
  const double dx = p->boxlen_x / (double)(p->nx - 1);
  const double dy = p->boxlen_y / (double)(p->ny - 1);
  const double r121 = (double)1.21f, r03 = (double)0.3f;
  const double k8 = (double)(100.0f * 1.21f);  /* 100*1.21 in default real */
  for (int i = 0; i < p->nx; ++i)
    for (int j = 0; j < p->ny; ++j) {
      const double x = (double)i * dx, y = (double)j * dy;
      double w[4];
      if (ninit == 1) {
        const double ax = x - p->boxlen_x / 2.0, ay = y - p->boxlen_y / 2.0;
        w[0] = 1.0 + exp(-(ax * ax + ay * ay) * 20.0);
        w[1] = 1.0;
        w[2] = 1.0;
        w[3] = 1.0;
      } else if (ninit == 2) {
        const double rho_0 = r121, p_0 = 1.0, g = 1.0;
        w[0] = rho_0 * exp(-(rho_0 * g / p_0) * (x + y));
        w[1] = 0.0;
        w[2] = 0.0;
        w[3] = p_0 * exp(-(rho_0 * g / p_0) * (x + y));
      } else if (ninit == 8) {
        const double bx = x - r03, by = y - r03;
        w[0] = r121 * exp(-r121 * (x + y));
        w[1] = 0.0;
        w[2] = 0.0;
        w[3] = exp(-r121 * (x + y)) + eta * exp(-(k8 * (bx * bx + by * by)));
      } else {
        return -1;
      }
      orc_rd_conservative(p->gamma, w, &U4(u, 0, i, j, p->nx));
    }
  return 0;
}
