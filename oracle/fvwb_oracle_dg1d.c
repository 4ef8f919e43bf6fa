/*
 * fvwb_oracle_dg1d.c — CPU FP64 restatement of the reference's 1D modal DG with
 * source (dg_with_source.f90 with legendre.f90 and dg_commons.f90).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker; the product never links it.
 *
 * Every function follows the reference statement by statement (built with
 * -ffp-contract=off), including its default-real literals: sqrt(0.5), sqrt(6.),
 * sqrt(10.), sqrt(3.)/sqrt(5.), 5./9., 8./9. and 0.125/n/n are FP32 values.
 * Integer powers x**k are evaluated by binary powering, x**3 = x*(x*x),
 * x**4 = (x*x)*(x*x), as the compiler's integer-power routine does.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "fvwb_oracle.h"

#define V3(a, v, m, c, n) ((a)[(size_t)(v) + 3u * ((size_t)(m) + (size_t)(n) * (size_t)(c))])

static double ipow(double a, int b)
{
  double r = 1;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return r;
}

/* legendre.f90:1-24 */
double orc_dg1d_legendre(double x, int n)
{
  x = fmin(fmax(x, -1.0), 1.0);
  switch (n) {
  case 0: return 1.0 * (double)sqrtf(0.5f);
  case 1: return x * 0.5 * (double)sqrtf(6.f);
  case 2: return 0.25 * (3.0 * ipow(x, 2) - 1.0) * (double)sqrtf(10.f);
  case 3: return 0.5 * (5.0 * ipow(x, 3) - 3.0 * x);
  case 4: return 0.125 * (35.0 * ipow(x, 4) - 30.0 * ipow(x, 2) + 3.0);
  case 5: return 0.125 * (63.0 * ipow(x, 5) - 70.0 * ipow(x, 3) + 15.0 * x);
  case 6: return 1.0 / 16.0 * (231.0 * ipow(x, 6) - 315.0 * ipow(x, 4) + 105.0 * ipow(x, 2) - 5.0);
  default: return NAN;
  }
}

/* legendre.f90:26-48; case 1 is the FP32 product 1.0*0.5*sqrt(6.) */
double orc_dg1d_legendre_prime(double x, int n)
{
  x = fmin(fmax(x, -1.0), 1.0);
  switch (n) {
  case 0: return 0.0;
  case 1: return (double)(1.0f * 0.5f * sqrtf(6.f));
  case 2: return 6.0 * x * 0.25 * (double)sqrtf(10.f);
  case 3: return 0.5 * (15.0 * ipow(x, 2) - 3.0);
  case 4: return 0.125 * (140.0 * ipow(x, 3) - 60.0 * x);
  case 5: return 0.125 * (315.0 * ipow(x, 4) - 210.0 * ipow(x, 2) + 15.0);
  case 6: return 1.0 / 16.0 * (1386.0 * ipow(x, 5) - 1260.0 * ipow(x, 3) + 210.0 * x);
  default: return NAN;
  }
}

/* legendre.f90:77-127: literals for n <= 3, else 50 Newton steps from the usual
 * guess, the weight 2(2n+1)/((1-x^2) P_n'(x)^2) as written, and the second half
 * mirrored from the first (roots in descending order, positive first). */
void orc_dg1d_gl_quadrature(int n, double* x, double* w)
{
  if (n == 1) {
    x[0] = 0.0; w[0] = 2.0;
    return;
  }
  if (n == 2) {
    x[0] = -(1. / (double)3 * (double)sqrtf(3.f));
    x[1] = 1. / (double)3 * (double)sqrtf(3.f);
    w[0] = 1.; w[1] = 1.;
    return;
  }
  if (n == 3) {
    x[0] = (double)(-(sqrtf(3.f) / sqrtf(5.f)));
    x[1] = 0.0;
    x[2] = (double)(sqrtf(3.f) / sqrtf(5.f));
    w[0] = (double)(5.f / 9.f); w[1] = (double)(8.f / 9.f); w[2] = (double)(5.f / 9.f);
    return;
  }
  const double dpi = acos(-1.0);
  const float nf = (float)n;
  const float a = 1.0f - 0.125f / nf / nf + 0.125f / nf / nf / nf;
  for (int i = 1; i <= n; ++i) {
    double xx = (double)a * cos(dpi * (4.0 * (double)i - 1.0) / (4.0 * (double)n + 2.0));
    for (int it = 0; it < 50; ++it) xx = xx - orc_dg1d_legendre(xx, n) / orc_dg1d_legendre_prime(xx, n);
    x[i - 1] = xx;
    const double lp = orc_dg1d_legendre_prime(x[i - 1], n);
    w[i - 1] = 2.0 * (2.0 * (double)n + 1.0) / (1.0 - ipow(x[i - 1], 2)) / ipow(lp, 2);
  }
  for (int i = n / 2 + 1; i <= n; ++i) {
    x[i - 1] = -x[n - i];
    w[i - 1] = w[n - i];
  }
}

/* compute_primitive / compute_conservative / compute_flux / compute_source of
 * dg_with_source.f90:1181-1246 (the same arithmetic as benchmark_1d.f90's) */
static void dg1d_source(double gamma, const double u[3], double s[3])
{
  double w[3];
  orc_ref1d_primitive(gamma, u, w);
  s[0] = 0;
  s[1] = -w[0];
  s[2] = -w[0] * w[1];
}

static void dg1d_riemann(const fvwb_dg1d_params* p, const double ul[3], const double ur[3], double f[3])
{
  if (p->riemann == FVWB_DG1D_RIEMANN_HLLC) {
    orc_ref_riemann_hllc_1d(p->gamma, ul, ur, f);
  } else {
    /* riemann_llf (:1299-1316) is compute_llflux of benchmark_1d.f90 with the
     * fluxes computed inside */
    double fl[3], fr[3];
    orc_ref1d_flux(p->gamma, ul, fl);
    orc_ref1d_flux(p->gamma, ur, fr);
    orc_ref1d_llflux(p->gamma, ul, ur, fl, fr, f);
  }
}

void orc_dg1d_tables(int n, double* chsi, double* wq, double* phi_q, double* dphi_q, double* phi_lr)
{
  orc_dg1d_gl_quadrature(n, chsi, wq);
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) {
      phi_q[i * n + j] = orc_dg1d_legendre(chsi[j], i);
      dphi_q[i * n + j] = orc_dg1d_legendre_prime(chsi[j], i);
    }
    phi_lr[2 * i] = orc_dg1d_legendre(-1.0, i);
    phi_lr[2 * i + 1] = orc_dg1d_legendre(1.0, i);
  }
}

/* compute_update_exact_delta (dg_with_source.f90:1749-2031) */
void orc_dg1d_update_exact_delta(const fvwb_dg1d_params* p, const double* du, const double* ueq, double* dudt)
{
  const int nx = p->nx, n = p->n;
  const double gamma = p->gamma;
  const double dx = p->boxlen / (double)nx;
  const double oneoverdx = 1. / dx;
  double chsi[8], wq[8], phq[64], dphq[64], plr[16];
  orc_dg1d_tables(n, chsi, wq, phq, dphq, plr);

  double* ufe = malloc(sizeof(double) * 3 * (nx + 1));    /* u_face_eq */
  double* ffe = malloc(sizeof(double) * 3 * (nx + 1));    /* flux_face_eq */
  double* ff = calloc(3 * (size_t)(nx + 1), sizeof(double)); /* flux_face */
  double* ul = malloc(sizeof(double) * 3 * nx);
  double* ur = malloc(sizeof(double) * 3 * nx);
  for (int i = 0; i < nx + 1; ++i) {
    const double xf = (double)i * dx;
    const double w[3] = {exp(-xf), 0, exp(-xf)};
    orc_ref1d_conservative(gamma, w, &ufe[3 * i]);
    orc_ref1d_flux(gamma, &ufe[3 * i], &ffe[3 * i]);
  }
  /* traces at xi = -1 and +1 plus the face equilibrium (:1851-1895) */
  for (int c = 0; c < nx; ++c) {
    double dl[3] = {0., 0., 0.}, dr[3] = {0., 0., 0.};
    for (int i = 0; i < n; ++i)
      for (int v = 0; v < 3; ++v) {
        dl[v] = dl[v] + V3(du, v, i, c, n) * plr[2 * i];
        dr[v] = dr[v] + V3(du, v, i, c, n) * plr[2 * i + 1];
      }
    for (int v = 0; v < 3; ++v) {
      ul[3 * c + v] = ufe[3 * c + v] + dl[v];
      ur[3 * c + v] = ufe[3 * (c + 1) + v] + dr[v];
    }
  }
  /* interior faces: zero-based face f between cells f-1 and f (the boundary
   * faces reach only the two end cells, whose dudt is zeroed) */
  for (int f = 1; f < nx; ++f) dg1d_riemann(p, &ur[3 * (f - 1)], &ul[3 * f], &ff[3 * f]);

  for (int c = 1; c < nx - 1; ++c) {
    double fq[8][3], fqe[8][3], sq[8][3], sqe[8][3];
    for (int j = 0; j < n; ++j) {
      double uq[3] = {0.0, 0.0, 0.0}, us[3];
      for (int i = 0; i < n; ++i)
        for (int v = 0; v < 3; ++v) uq[v] = uq[v] + V3(du, v, i, c, n) * phq[i * n + j];
      for (int v = 0; v < 3; ++v) us[v] = V3(ueq, v, j, c, n) + uq[v];
      orc_ref1d_flux(gamma, us, fq[j]);
      orc_ref1d_flux(gamma, &V3(ueq, 0, j, c, n), fqe[j]);
      if (p->source == 2) {
        dg1d_source(gamma, us, sq[j]);
        dg1d_source(gamma, &V3(ueq, 0, j, c, n), sqe[j]);
      }
    }
    for (int i = 0; i < n; ++i) {
      double fv[3] = {0.0, 0.0, 0.0}, fve[3] = {0.0, 0.0, 0.0};
      double sv[3] = {0.0, 0.0, 0.0}, sve[3] = {0.0, 0.0, 0.0};
      for (int j = 0; j < n; ++j)
        for (int v = 0; v < 3; ++v) {
          fv[v] = fv[v] + fq[j][v] * dphq[i * n + j] * wq[j];
          fve[v] = fve[v] + fqe[j][v] * dphq[i * n + j] * wq[j];
          if (p->source == 2) {
            sv[v] = sv[v] + sq[j][v] * phq[i * n + j] * wq[j] * 0.5;
            sve[v] = sve[v] + sqe[j][v] * phq[i * n + j] * wq[j] * 0.5;
          }
        }
      const double pl = plr[2 * i], pr = plr[2 * i + 1];
      for (int v = 0; v < 3; ++v)
        V3(dudt, v, i, c, n) = oneoverdx * fv[v] - oneoverdx * fve[v] -
                               oneoverdx * (ff[3 * (c + 1) + v] * pr - ff[3 * c + v] * pl) +
                               oneoverdx * (ffe[3 * (c + 1) + v] * pr - ffe[3 * c + v] * pl) + sv[v] - sve[v];
    }
  }
  for (int i = 0; i < n; ++i)
    for (int v = 0; v < 3; ++v) {
      V3(dudt, v, i, 0, n) = 0;
      V3(dudt, v, i, nx - 1, n) = 0;
    }
  free(ufe); free(ffe); free(ff); free(ul); free(ur);
}

/* the SSP(5,4) coefficients of the 'RKi' branch (dg_with_source.f90:264-288) as
 * its default-real literals round them, in the order
 * c10; c20 c21 c22; c30 c32 c33; c40 c43 c44; c52 c53 c54; c55 c56 */
void orc_dg1d_rki_coeffs(double c[15])
{
  static const float k[15] = {0.391752226571890f, 0.444370493651235f, 0.555629506348765f, 0.368410593050371f,
                              0.620101851488403f, 0.379898148511597f, 0.251891774271694f, 0.178079954393132f,
                              0.821920045606868f, 0.544974750228521f, 0.517231671970585f, 0.096059710526147f,
                              0.063692468666290f, 0.386708617503269f, 0.226007483236906f};
  for (int i = 0; i < 15; ++i) c[i] = (double)k[i];
}

/* The main program's set-up (dg_with_source.f90:27-31, 105-125, 148-167) for
 * condinit cases 7 (rest) and 8 (rest plus a pressure pulse of size pert at
 * boxlen/2): u_eq and u_init at each cell's Gauss points xcell + dx/2*chsi_j,
 * and the initial modes sum_j 0.5*(u_init - u_eq)*phi_i(chsi_j)*w_j. */
void orc_dg1d_initial(const fvwb_dg1d_params* p, int ninit, double pert, double* du, double* ueq, double* uinit)
{
  const int nx = p->nx, n = p->n;
  const double gamma = p->gamma;
  const double dx = p->boxlen / (double)nx;
  double chsi[8], wq[8], phq[64], dphq[64], plr[16];
  orc_dg1d_tables(n, chsi, wq, phq, dphq, plr);
  for (int c = 0; c < nx; ++c) {
    const double xcell = ((double)(c + 1) - 0.5) * dx;
    for (int j = 0; j < n; ++j) {
      const double x = xcell + dx / 2.0 * chsi[j];
      const double we[3] = {exp(-x), 0, exp(-x)};
      double wi[3] = {exp(-x), 0., exp(-x)};
      if (ninit == 8) wi[2] = exp(-x) + pert * exp(-(100 * ipow(x - p->boxlen / 2., 2)));
      orc_ref1d_conservative(gamma, we, &V3(ueq, 0, j, c, n));
      orc_ref1d_conservative(gamma, wi, &V3(uinit, 0, j, c, n));
    }
    for (int i = 0; i < n; ++i)
      for (int v = 0; v < 3; ++v) {
        double a = 0.0;
        for (int j = 0; j < n; ++j)
          a = a + 0.5 * (V3(uinit, v, j, c, n) - V3(ueq, v, j, c, n)) * phq[i * n + j] * wq[j];
        V3(du, v, i, c, n) = a;
      }
  }
}

/* The main loop with integrator 'RKi' (dg_with_source.f90:191-199, 264-288,
 * 303-334): dt = 0.9*dx/cmax/(2n+1) with cmax over cells of |v|+c at the first
 * Gauss point of uinit; the five stages on the modes; then uinit = u_eq plus
 * the modes evaluated at legendre(x_quad) with x_quad the point's PHYSICAL
 * coordinate xcell + dx/2*chsi_i, as written (:314-323); t = t + dt.  Runs
 * while t < t_end and fewer than max_steps steps were taken; writes each
 * step's t and dt to ts and dts when given; uinit is updated in place. */
int orc_dg1d_run_rki(const fvwb_dg1d_params* p, double* du, const double* ueq, double* uinit, double t_end,
                     int max_steps, double* t_out, double* ts, double* dts)
{
  const int nx = p->nx, n = p->n;
  const size_t m = 3 * (size_t)n * (size_t)nx;
  const double dx = p->boxlen / (double)nx;
  double c[15], chsi[8], wq[8], phq[64], dphq[64], plr[16];
  orc_dg1d_rki_coeffs(c);
  orc_dg1d_tables(n, chsi, wq, phq, dphq, plr);
  double* L = malloc(sizeof(double) * m);
  double* w1 = malloc(sizeof(double) * m);
  double* w2 = malloc(sizeof(double) * m);
  double* w3 = malloc(sizeof(double) * m);
  double* w4 = malloc(sizeof(double) * m);
  double t = 0;
  int it = 0;
  while (t < t_end && it < max_steps) {
    double cmax = 0.0;
    for (int k = 0; k < nx; ++k) cmax = fmax(cmax, orc_ref1d_speed(p->gamma, &V3(uinit, 0, 0, k, n)));
    const double dt = (double)0.9f * dx / cmax / (2.0 * (double)n + 1.0);
    orc_dg1d_update_exact_delta(p, du, ueq, L);
    for (size_t k = 0; k < m; ++k) w1[k] = du[k] + c[0] * dt * L[k];
    orc_dg1d_update_exact_delta(p, w1, ueq, L);
    for (size_t k = 0; k < m; ++k) w2[k] = c[1] * du[k] + c[2] * w1[k] + c[3] * dt * L[k];
    orc_dg1d_update_exact_delta(p, w2, ueq, L);
    for (size_t k = 0; k < m; ++k) w3[k] = c[4] * du[k] + c[5] * w2[k] + c[6] * dt * L[k];
    orc_dg1d_update_exact_delta(p, w3, ueq, L);
    for (size_t k = 0; k < m; ++k) {
      w4[k] = c[7] * du[k] + c[8] * w3[k] + c[9] * dt * L[k];
      du[k] = c[10] * w2[k] + c[11] * w3[k] + c[12] * dt * L[k];
    }
    orc_dg1d_update_exact_delta(p, w4, ueq, L);
    for (size_t k = 0; k < m; ++k) du[k] = du[k] + c[13] * w4[k] + c[14] * dt * L[k];
    for (int k = 0; k < nx; ++k) {
      const double xcell = ((double)(k + 1) - 0.5) * dx;
      for (int i = 0; i < n; ++i) {
        const double xq = xcell + dx / 2.0 * chsi[i];
        for (int v = 0; v < 3; ++v) {
          double a = 0.0;
          for (int q = 0; q < n; ++q) a = a + V3(du, v, q, k, n) * orc_dg1d_legendre(xq, q);
          V3(uinit, v, i, k, n) = V3(ueq, v, i, k, n) + a;
        }
      }
    }
    t = t + dt;
    if (ts) ts[it] = t;
    if (dts) dts[it] = dt;
    ++it;
  }
  free(L); free(w1); free(w2); free(w3); free(w4);
  if (t_out) *t_out = t;
  return it;
}
