/*
 * fvwb_oracle_dgq.c — CPU FP64 oracle, the reference's own 2D modal DG program
 * (TEST INFRASTRUCTURE).
 *
 * Restates 2d/benchmark_2d_dg.f90 as the reference ships it, statement by
 * statement (Fortran expression order, no FMA: built with -ffp-contract=off,
 * default-real literals rounded to FP32 where the reference writes them):
 *   compute_update        :1137-1479  tensor (Q) Legendre basis of mx = my =
 *                                     order modes per direction, not well
 *                                     balanced (u_eq unused), flux dispatch
 *                                     :991-1006, sources :1558-1644, assembly
 *                                     :1448-1466, special_boundary_conditions
 *                                     :1481-1514
 *   compute_max_speed     :826-889    with its cs_max / v_xmax bookkeeping
 *   evolve                :624-775    dt rule :671, solvers 'RK4', 'SS4',
 *                                     'EQL', 'DEB', apply_limiter 'ONP'
 *   compute_error         :23-89      L_inf, Gauss-weighted L1 and L2
 * Pinned bit for bit against flang builds of the reference (oracle/_ref,
 * tests/golden/dgq_*.npz): the shipped set-up and variants of its compile-time
 * switches.  Not linked by the product.
 *
 * Layout: modes and nodal values AoS(4, nx, ny, o, o), element (v,i,j,a,b) at
 * v + 4*(i + nx*(j + ny*(a + o*b))); coordinates (nx, ny, o, o).
 */
#include "fvwb_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define DQ(arr, v, i, j, a, b) ((arr)[(size_t)(v) + 4u * ((size_t)(i) + nx * ((size_t)(j) + ny * ((size_t)(a) + o * (size_t)(b))))])
#define XQ(arr, i, j, a, b) ((arr)[(size_t)(i) + nx * ((size_t)(j) + ny * ((size_t)(a) + o * (size_t)(b)))])

/* the reference's default-real literal 10e-10 (compute_primitive's density floor) */
static const double RHO_FLOOR = (double)10e-10f;

/* compute_primitive :891-902 for one point */
static void prim(double gamma, const double u[4], double w[4])
{
  w[0] = u[0] > RHO_FLOOR ? u[0] : RHO_FLOOR;
  w[1] = u[1] / w[0];
  w[2] = u[2] / w[0];
  w[3] = (gamma - 1.0) * (u[3] - 0.5 * w[0] * (w[1] * w[1] + w[2] * w[2]));
}

/* compute_flux :919-944 for one point: f1 along x, f2 along y */
static void flux_full(double gamma, const double u[4], double f1[4], double f2[4])
{
  double w[4];
  prim(gamma, u, w);
  f2[0] = w[0] * w[2];
  f2[1] = w[0] * w[1] * w[2];
  f2[2] = w[2] * u[2] + w[3];
  f2[3] = w[2] * u[3] + w[2] * w[3];
  f1[0] = w[0] * w[1];
  f1[1] = w[1] * u[1] + w[3];
  f1[2] = w[0] * w[1] * w[2];
  f1[3] = w[1] * u[3] + w[1] * w[3];
}

/* compute_flux_int :946-965: the face-trace flux, mass flux from u(1) */
static void flux_int(double gamma, const double u[4], double f1[4], double f2[4])
{
  double w[4];
  prim(gamma, u, w);
  f1[0] = w[1] * u[0];
  f1[1] = w[1] * u[1] + w[3];
  f1[2] = w[0] * w[1] * w[2];
  f1[3] = w[1] * u[3] + w[1] * w[3];
  f2[0] = w[2] * u[0];
  f2[1] = w[0] * w[1] * w[2];
  f2[2] = w[2] * u[2] + w[3];
  f2[3] = w[2] * u[3] + w[2] * w[3];
}

/* compute_speed :873-889 */
static void speed(double gamma, const double u[4], double* cs, double* vx, double* vy, double* sp)
{
  double w[4];
  prim(gamma, u, w);
  *cs = sqrt(gamma * fmax(w[3], 1e-10) / fmax(w[0], 1e-10));
  *vx = w[1];
  *vy = w[2];
  *sp = sqrt(w[1] * w[1] + w[2] * w[2]) + *cs;
}

static double fmax_f(double a, double b) { return a > b ? a : b; }
static double fmin_f(double a, double b) { return a < b ? a : b; }

/* compute_llflux :968-988 */
static void llflux(double gamma, const double ul[4], const double ur[4], const double fl[4], const double fr[4],
                   double out[4], int flag)
{
  double csl, vxl, vyl, spl, csr, vxr, vyr, spr;
  speed(gamma, ul, &csl, &vxl, &vyl, &spl);
  speed(gamma, ur, &csr, &vxr, &vyr, &spr);
  double cmax = 0.0;
  if (flag == 1) cmax = fmax_f(fabs(vxr + csr), fabs(vxl + csl));
  else if (flag == 2) cmax = fmax_f(fabs(vyr + csr), fabs(vyl + csl));
  for (int v = 0; v < 4; ++v) out[v] = 0.5 * (fr[v] + fl[v]) + 0.5 * cmax * (ul[v] - ur[v]);
}

/* compute_hllflux :1008-1026 (|v| in both speed bounds, no flag) */
static void hllflux(double gamma, const double ul[4], const double ur[4], const double fl[4], const double fr[4],
                    double out[4])
{
  double csl, vxl, vyl, spl, csr, vxr, vyr, spr;
  speed(gamma, ul, &csl, &vxl, &vyl, &spl);
  speed(gamma, ur, &csr, &vxr, &vyr, &spr);
  const double ap = fmax_f(0.0, fmax_f(csl + sqrt(vxl * vxl + vyl * vyl), csr + sqrt(vxr * vxr + vyr * vyr)));
  const double am = fmax_f(0.0, fmax_f(-(csl - sqrt(vxl * vxl + vyl * vyl)), -(csr - sqrt(vxr * vxr + vyr * vyr))));
  for (int v = 0; v < 4; ++v) out[v] = (ap * fl[v] + am * fr[v] - ap * am * (ur[v] - ul[v])) / (ap + am);
}

/* compute_hllcflux :1030-1134 as written: the star states use the left
 * tangential velocity in both y-direction star momenta and the right star
 * energy's parentheses as the reference places them; the branch fluxes come
 * from compute_flux (mass flux rho*v).  A state that satisfies no branch (a
 * NaN speed) leaves the output untouched, as the reference's does. */
static void hllcflux(double gamma, const double ul[4], const double ur[4], double out[4], int flag)
{
  double wl[4], wr[4], csl, vxl, vyl, spl, csr, vxr, vyr, spr;
  prim(gamma, ul, wl);
  prim(gamma, ur, wr);
  speed(gamma, ul, &csl, &vxl, &vyl, &spl);
  speed(gamma, ur, &csr, &vxr, &vyr, &spr);
  const double vl = flag == 1 ? vxl : vyl, vr = flag == 1 ? vxr : vyr;
  const double SL = fmin_f(vl, vr) - fmax_f(csl, csr);
  const double SR = fmax_f(vl, vr) + fmax_f(csl, csr);
  const double SM = (wr[0] * vr * (SR - vr) - wl[0] * vl * (SL - vl) + wl[3] - wr[3]) /
                    (wr[0] * (SR - vr) - wl[0] * (SL - vl));
  double usl[4], usr[4];
  usl[0] = ul[0] * (SL - vl) / (SL - SM);
  usr[0] = ur[0] * (SR - vr) / (SR - SM);
  if (flag == 1) {
    usl[1] = usl[0] * SM;
    usl[2] = usl[0] * wl[2];
    usl[3] = usl[0] * (ul[3] / ul[0] + (SM - wl[1]) * (SM + wl[3] / (wl[0] * (SL - wl[1]))));
    usr[1] = usr[0] * SM;
    usr[2] = usr[0] * wr[2];
    usr[3] = usr[0] * (ur[3] / ur[0] + (SM - wr[1] * (SM + wr[3] / (wr[0] * (SR - wr[1])))));
  } else {
    usl[1] = usl[0] * wl[1];
    usl[2] = usl[0] * SM;
    usl[3] = usl[0] * (ul[3] / ul[0] + (SM - wl[2]) * (SM + wl[3] / (wl[0] * (SL - wl[2]))));
    usr[1] = usr[0] * wl[1];
    usr[2] = usr[0] * SM;
    usr[3] = usr[0] * (ur[3] / ur[0] + (SM - wr[2] * (SM + wr[3] / (wr[0] * (SR - wr[2])))));
  }
  double f1[4], f2[4];
  const double* f = flag == 1 ? f1 : f2;
  if (SL > 0.0) {
    flux_full(gamma, ul, f1, f2);
    for (int v = 0; v < 4; ++v) out[v] = f[v];
  } else if (SL <= 0.0 && SM > 0.0) {
    flux_full(gamma, ul, f1, f2);
    for (int v = 0; v < 4; ++v) out[v] = f[v] + SL * (usl[v] - ul[v]);
  } else if (SR >= 0.0 && SM <= 0.0) {
    flux_full(gamma, ur, f1, f2);
    for (int v = 0; v < 4; ++v) out[v] = f[v] + SR * (usr[v] - ur[v]);
  } else if (SR < 0.0) {
    flux_full(gamma, ur, f1, f2);
    for (int v = 0; v < 4; ++v) out[v] = f[v];
  }
}

/* compute_num_flux :991-1006: 'llf1', 'hll2', 'hllc'; any other flux_type
 * (the shipped 'llf') matches no branch and leaves the face flux at zero */
static void num_flux(const fvwb_dgq_params* p, const double ul[4], const double ur[4], const double fl[4],
                     const double fr[4], double out[4], int flag)
{
  switch (p->flux_type) {
  case FVWB_DGQ_FLUX_LLF1: llflux(p->gamma, ul, ur, fl, fr, out, flag); break;
  case FVWB_DGQ_FLUX_HLL2: hllflux(p->gamma, ul, ur, fl, fr, out); break;
  case FVWB_DGQ_FLUX_HLLC: hllcflux(p->gamma, ul, ur, out, flag); break;
  default: break;
  }
}

/* get_boundary_conditions(index, 2) :777-824 — the reference passes dim = 2
 * for the x faces too, so both directions wrap with ny (nx == ny) */
static int bc_index(int index, int n, int bc)
{
  if (bc == 1) {
    if (index == 0) return n;
    if (index == n + 1) return 1;
  } else if (bc == 2 || bc == 3) {
    if (index == 0) return 1;
    if (index == n + 1) return n;
  }
  return index;
}

/* grad_phi :1599-1644 at one point */
static void grad_phi(int gcase, double x, double y, double g[2])
{
  if (gcase == 1) {                      /* "linear phi = x+y": grad_p = (x, y) */
    g[0] = x;
    g[1] = y;
  } else if (gcase == 2) {               /* point mass at (3, 3), softened inside r = 0.45 */
    const double epsilon = 0.25, delta_r = (double)0.1f, xc = 3., yc = 3.;
    const double xd = x - xc, yd = y - yc;
    const double r = sqrt(xd * xd + yd * yd);
    if (r > 0.5 - 0.5 * delta_r) {
      const double r3 = r * r * r;
      g[0] = -(xd) / r3;
      g[1] = -(yd) / r3;
    } else {
      g[0] = -(xd) / (r * (r * r + epsilon * epsilon));
      g[1] = -(yd) / (r * (r * r + epsilon * epsilon));
    }
  } else {
    g[0] = g[1] = 0.0;
  }
}

int orc_refdg_valid(const fvwb_dgq_params* p)
{
  return p->nx >= 1 && p->nx == p->ny && p->order >= 2 && p->order <= 3;
}

void orc_refdg_update(const fvwb_dgq_params* p, const double* du, const double* x, const double* y, double* dudt)
{
  const size_t nx = (size_t)p->nx, ny = (size_t)p->ny, o = (size_t)p->order;
  const size_t npt = 4 * nx * ny * o * o;
  const double gamma = p->gamma;
  const double dx = p->boxlen_x / (double)nx;
  const double odx = 1. / dx;
  double xq[8], wq[8], lm[3], lp[3], Lq[3][3], Pq[3][3];
  orc_dg_gl_quadrature((int)o, xq, wq);
  for (size_t n = 0; n < o; ++n) {
    lm[n] = orc_dg_legendre(-1.0, (int)n);
    lp[n] = orc_dg_legendre(1.0, (int)n);
    for (size_t a = 0; a < o; ++a) {
      Lq[a][n] = orc_dg_legendre(xq[a], (int)n);
      Pq[a][n] = orc_dg_legendre_prime(xq[a], (int)n);
    }
  }
  double* nodes = malloc(sizeof(double) * npt);
  double* fq1 = malloc(sizeof(double) * npt);
  double* fq2 = malloc(sizeof(double) * npt);
  double* fv1 = calloc(npt, sizeof(double));
  double* fv2 = calloc(npt, sizeof(double));
  /* traces: [v + 4*(q + o*(i + nx*j))] */
  const size_t ntr = 4 * o * nx * ny;
  double* uL = malloc(sizeof(double) * ntr);
  double* uR = malloc(sizeof(double) * ntr);
  double* uB = malloc(sizeof(double) * ntr);
  double* uT = malloc(sizeof(double) * ntr);
  double* fL = malloc(sizeof(double) * 2 * ntr);   /* flux_left(:,...,dir) */
  double* fR = malloc(sizeof(double) * 2 * ntr);
  double* fB = malloc(sizeof(double) * 2 * ntr);
  double* fT = malloc(sizeof(double) * 2 * ntr);
  /* F(v, q, iface, j), iface 0..nx; G(v, q, i, jface), jface 0..ny */
  double* F = calloc(4 * o * (nx + 1) * ny, sizeof(double));
  double* G = calloc(4 * o * nx * (ny + 1), sizeof(double));
  double* s = calloc(npt, sizeof(double));
#define TR(arr, v, q, i, j) ((arr)[(size_t)(v) + 4 * ((size_t)(q) + o * ((size_t)(i) + nx * (size_t)(j)))])
#define FF(v, q, f, j) (F[(size_t)(v) + 4 * ((size_t)(q) + o * ((size_t)(f) + (nx + 1) * (size_t)(j)))])
#define GG(v, q, i, f) (G[(size_t)(v) + 4 * ((size_t)(q) + o * ((size_t)(i) + nx * (size_t)(f)))])

  /* :1201-1244 nodes, fluxes at the nodes, volume integrals */
  orc_ref_dg2d_nodes_from_modes((int)nx, (int)ny, (int)o, du, nodes);
  for (size_t k = 0; k < nx * ny * o * o; ++k) flux_full(gamma, &nodes[4 * k], &fq1[4 * k], &fq2[4 * k]);
  for (size_t ic = 0; ic < nx; ++ic)
    for (size_t jc = 0; jc < ny; ++jc)
      for (size_t i = 0; i < o; ++i)
        for (size_t j = 0; j < o; ++j) {
          for (int v = 0; v < 4; ++v) DQ(fv1, v, ic, jc, i, j) = 0.0;
          for (size_t a = 0; a < o; ++a)
            for (size_t b = 0; b < o; ++b)
              for (int v = 0; v < 4; ++v)
                DQ(fv1, v, ic, jc, i, j) = DQ(fv1, v, ic, jc, i, j) +
                                           DQ(fq1, v, ic, jc, a, b) * Pq[a][i] * wq[a] * Lq[b][j] * wq[b];
          for (size_t a = 0; a < o; ++a)
            for (size_t b = 0; b < o; ++b)
              for (int v = 0; v < 4; ++v)
                DQ(fv2, v, ic, jc, i, j) = DQ(fv2, v, ic, jc, i, j) +
                                           DQ(fq2, v, ic, jc, a, b) * Pq[b][j] * wq[b] * Lq[a][i] * wq[a];
        }

  /* :1253-1314 traces at xi = -1, +1 and eta = -1, +1 */
  for (size_t ic = 0; ic < nx; ++ic)
    for (size_t jc = 0; jc < ny; ++jc) {
      double l[4][3] = {{0}}, r[4][3] = {{0}};
      for (size_t i = 0; i < o; ++i)
        for (size_t j = 0; j < o; ++j)
          for (size_t q = 0; q < o; ++q)
            for (int v = 0; v < 4; ++v) {
              l[v][q] = l[v][q] + DQ(du, v, ic, jc, i, j) * lm[i] * Lq[q][j];
              r[v][q] = r[v][q] + DQ(du, v, ic, jc, i, j) * lp[i] * Lq[q][j];
            }
      for (size_t q = 0; q < o; ++q)
        for (int v = 0; v < 4; ++v) {
          TR(uL, v, q, ic, jc) = l[v][q];
          TR(uR, v, q, ic, jc) = r[v][q];
        }
    }
  for (size_t ic = 0; ic < nx; ++ic)
    for (size_t jc = 0; jc < ny; ++jc) {
      double bt[4][3] = {{0}}, tp[4][3] = {{0}};
      for (size_t i = 0; i < o; ++i)
        for (size_t j = 0; j < o; ++j)
          for (size_t q = 0; q < o; ++q)
            for (int v = 0; v < 4; ++v) {
              bt[v][q] = bt[v][q] + DQ(du, v, ic, jc, i, j) * lm[j] * Lq[q][i];
              tp[v][q] = tp[v][q] + DQ(du, v, ic, jc, i, j) * lp[j] * Lq[q][i];
            }
      for (size_t q = 0; q < o; ++q)
        for (int v = 0; v < 4; ++v) {
          TR(uB, v, q, ic, jc) = bt[v][q];
          TR(uT, v, q, ic, jc) = tp[v][q];
        }
    }
  for (size_t ic = 0; ic < nx; ++ic)
    for (size_t jc = 0; jc < ny; ++jc)
      for (size_t q = 0; q < o; ++q) {
        const size_t k = 4 * (q + o * (ic + nx * jc));
        flux_int(gamma, &uL[k], &fL[2 * k], &fL[2 * k + 4]);
        flux_int(gamma, &uR[k], &fR[2 * k], &fR[2 * k + 4]);
        flux_int(gamma, &uT[k], &fT[2 * k], &fT[2 * k + 4]);
        flux_int(gamma, &uB[k], &fB[2 * k], &fB[2 * k + 4]);
      }

  /* :1317-1366 face fluxes */
  for (size_t j = 0; j < ny; ++j)
    for (size_t f = 0; f <= nx; ++f) {
      const int il = bc_index((int)f, (int)ny, p->bc) - 1, ir = bc_index((int)f + 1, (int)ny, p->bc) - 1;
      for (size_t q = 0; q < o; ++q) {
        const size_t kl = 4 * (q + o * ((size_t)il + nx * j)), kr = 4 * (q + o * ((size_t)ir + nx * j));
        num_flux(p, &uR[kl], &uL[kr], &fR[2 * kl], &fL[2 * kr], &FF(0, q, f, j), 1);
      }
    }
  for (size_t i = 0; i < nx; ++i)
    for (size_t f = 0; f <= ny; ++f) {
      const int il = bc_index((int)f, (int)ny, p->bc) - 1, ir = bc_index((int)f + 1, (int)ny, p->bc) - 1;
      for (size_t q = 0; q < o; ++q) {
        const size_t kl = 4 * (q + o * (i + nx * (size_t)il)), kr = 4 * (q + o * (i + nx * (size_t)ir));
        num_flux(p, &uT[kl], &uB[kr], &fT[2 * kl + 4], &fB[2 * kr + 4], &GG(0, q, i, f), 2);
      }
    }

  /* :1372-1411 edge integrals, :1413-1443 source, :1448-1466 assembly */
  if (p->source == 2) {
    for (size_t ic = 0; ic < nx; ++ic)
      for (size_t jc = 0; jc < ny; ++jc)
        for (size_t a = 0; a < o; ++a)
          for (size_t b = 0; b < o; ++b) {
            double w[4], g[2], u[4];
            for (int v = 0; v < 4; ++v) u[v] = DQ(nodes, v, ic, jc, a, b);
            prim(gamma, u, w);
            grad_phi(p->grad_phi_case, XQ(x, ic, jc, a, b), XQ(y, ic, jc, a, b), g);
            DQ(s, 0, ic, jc, a, b) = 0.;
            DQ(s, 1, ic, jc, a, b) = w[0] * g[0];
            DQ(s, 2, ic, jc, a, b) = w[0] * g[1];
            DQ(s, 3, ic, jc, a, b) = w[0] * (w[1] * g[0] + w[2] * g[1]);
          }
  } else if (p->source == 3) {
    for (size_t ic = 0; ic < nx; ++ic)
      for (size_t jc = 0; jc < ny; ++jc)
        for (size_t a = 0; a < o; ++a)
          for (size_t b = 0; b < o; ++b) DQ(s, 0, ic, jc, a, b) = -1.0 * DQ(nodes, 0, ic, jc, a, b);
  }
  for (size_t ic = 0; ic < nx; ++ic)
    for (size_t jc = 0; jc < ny; ++jc)
      for (size_t i = 0; i < o; ++i)
        for (size_t j = 0; j < o; ++j) {
          double e1[4] = {0}, e2[4] = {0}, e3[4] = {0}, e4[4] = {0}, sv[4] = {0};
          for (size_t a = 0; a < o; ++a)
            for (int v = 0; v < 4; ++v) {
              e1[v] = e1[v] + FF(v, a, ic + 1, jc) * lp[i] * Lq[a][j] * wq[a];
              e2[v] = e2[v] + FF(v, a, ic, jc) * lm[i] * Lq[a][j] * wq[a];
            }
          for (size_t a = 0; a < o; ++a)
            for (int v = 0; v < 4; ++v) {
              e3[v] = e3[v] + GG(v, a, ic, jc + 1) * lp[j] * Lq[a][i] * wq[a];
              e4[v] = e4[v] + GG(v, a, ic, jc) * lm[j] * Lq[a][i] * wq[a];
            }
          for (size_t a = 0; a < o; ++a)
            for (size_t b = 0; b < o; ++b)
              for (int v = 0; v < 4; ++v)
                sv[v] = sv[v] + DQ(s, v, ic, jc, a, b) * Lq[a][i] * wq[a] * Lq[b][j] * wq[b];
          for (int v = 0; v < 4; ++v)
            DQ(dudt, v, ic, jc, i, j) = (odx * DQ(fv1, v, ic, jc, i, j) + odx * DQ(fv2, v, ic, jc, i, j) -
                                         odx * (e1[v] - e2[v]) - odx * (e3[v] - e4[v])) / 2. +
                                        sv[v] / 4.;
        }
  if (p->ninit == 12) {                /* special_boundary_conditions :1481-1514 */
    const double xc = p->boxlen_x / 2., yc = p->boxlen_y / 2.;
    for (size_t ic = 0; ic < nx; ++ic)
      for (size_t jc = 0; jc < ny; ++jc)
        for (size_t a = 0; a < o; ++a)
          for (size_t b = 0; b < o; ++b) {
            const double xd = XQ(x, ic, jc, a, b) - xc, yd = XQ(y, ic, jc, a, b) - yc;
            if (sqrt(xd * xd + yd * yd) > 2.0)
              for (int v = 0; v < 4; ++v) DQ(dudt, v, ic, jc, a, b) = 0.0;
          }
  }
#undef TR
#undef FF
#undef GG
  free(nodes); free(fq1); free(fq2); free(fv1); free(fv2);
  free(uL); free(uR); free(uB); free(uT); free(fL); free(fR); free(fB); free(fT);
  free(F); free(G); free(s);
}

/* compute_max_speed :826-870 on u(1:4,1:nx,1:ny) (evolve passes the first
 * mode, the cell means): the last cell, in the reference's x-outer order,
 * whose speed reaches the running maximum gives v_xmax and v_ymax; cs_max
 * restarts at that cell's sound speed and then keeps the minimum of every
 * later cell's.  out = (cs_max, v_xmax, v_ymax, speed_max). */
void orc_refdg_max_speed(const fvwb_dgq_params* p, const double* u, double out[4])
{
  double smax = 0.0, csmax = 0.0, vxm = 0.0, vym = 0.0;
  for (int i = 0; i < p->nx; ++i)
    for (int j = 0; j < p->ny; ++j) {
      double cs, vx, vy, sp;
      speed(p->gamma, &u[4 * ((size_t)i + (size_t)p->nx * j)], &cs, &vx, &vy, &sp);
      if (sp >= smax) {
        smax = fmax_f(smax, sp);
        vxm = vx;
        vym = vy;
        csmax = cs;
      }
      if (csmax > cs) csmax = cs;
    }
  out[0] = csmax; out[1] = vxm; out[2] = vym; out[3] = smax;
}

/* evolve :668-671: gll_w_1 (in default REAL: 1e-10 vanishes against
 * gll(gll-1)) and the time step from compute_max_speed of the means; dy is
 * never used (both denominators take dx) */
double orc_refdg_dt(const fvwb_dgq_params* p, const double* modes, double t, double t_end, double ms[4])
{
  const int o = p->order, gll = (2 * (o - 1) + 3) / 2;
  const double gll_w_1 = o == 1 ? 1. : 1. / (double)((float)(gll * (gll - 1)) + 1e-10f);
  const double dx = p->boxlen_x / (double)p->nx;
  const size_t nc = (size_t)p->nx * p->ny;
  double* mean = malloc(sizeof(double) * 4 * nc);
  memcpy(mean, modes, sizeof(double) * 4 * nc);            /* delta_u(:,:,:,1,1) */
  orc_refdg_max_speed(p, mean, ms);
  free(mean);
  const double c1 = 1. / (double)(2 * 4 + 1), c2 = gll_w_1 / 2.;
  const double dt = p->cfl * fmin_f(c1, c2) / ((fabs(ms[1]) + ms[0]) / dx + (fabs(ms[2]) + ms[0]) / dx);
  return fmin_f(t_end - t, dt);
}

void orc_refdg_limiter(const fvwb_dgq_params* p, double* u)
{
  if (p->use_limiter) orc_ref_dg2d_positivity(p->nx, p->ny, p->order, p->gamma, p->eps, u);
}

/* evolve :666-769 from a limited state: at most max_steps steps while
 * t < t_end; dt of every step into dts (if not NULL); returns the steps taken */
int orc_refdg_evolve(const fvwb_dgq_params* p, double* du, const double* x, const double* y, double t_end,
                     int max_steps, double* t_io, double* dts)
{
  const size_t n = 4u * (size_t)p->nx * p->ny * p->order * p->order;
  double *L = malloc(sizeof(double) * n), *w1 = malloc(sizeof(double) * n), *w2 = malloc(sizeof(double) * n),
         *w3 = malloc(sizeof(double) * n), *w4 = malloc(sizeof(double) * n), *w5 = malloc(sizeof(double) * n);
  /* the 'RK4' (:683-710) and 'SS4' (:711-735) literals, default REAL */
  static const float rk4[16] = {0.391752226571890f, 0.444370493651235f, 0.555629506348765f, 0.368410593050371f,
                                0.620101851488403f, 0.379898148511597f, 0.251891774271694f, 0.178079954393132f,
                                0.821920045606868f, 0.544974750228521f, 0.00683325884039f, 0.51723167208978f,
                                0.12759831133288f,  0.34833675773694f,  0.08460416338212f,  0.22600748319395f};
  static const float ss4[16] = {0.39175222700392f, 0.44437049406734f, 0.55562950593266f, 0.36841059262959f,
                                0.6201018513854f,  0.37989814861460f, 0.25189177424738f, 0.17807995410773f,
                                0.82192004589227f, 0.54497475021237f, 0.00683325884039f, 0.51723167208978f,
                                0.12759831133288f, 0.34833675773694f, 0.08460416338212f, 0.22600748319395f};
  const float* cf = p->solver == FVWB_DGQ_SOLVER_SS4 ? ss4 : rk4;
  double c[16];
  for (int k = 0; k < 16; ++k) c[k] = (double)cf[k];
  double t = *t_io;
  int iter = 0;
  while (t < t_end && iter < max_steps) {
    double ms[4];
    const double dt = orc_refdg_dt(p, du, t, t_end, ms);
    if (p->solver == FVWB_DGQ_SOLVER_EQL) {
      orc_refdg_update(p, du, x, y, L);
      for (size_t q = 0; q < n; ++q) w1[q] = du[q] + dt * L[q];
      orc_refdg_limiter(p, w1);
      orc_refdg_update(p, w1, x, y, L);
      for (size_t q = 0; q < n; ++q) du[q] = 0.5 * du[q] + 0.5 * w1[q] + 0.5 * dt * L[q];
      orc_refdg_limiter(p, du);
    } else if (p->solver == FVWB_DGQ_SOLVER_DEB) {
      orc_refdg_update(p, du, x, y, L);
      for (size_t q = 0; q < n; ++q) w1[q] = du[q] + dt * L[q];
      orc_refdg_limiter(p, w1);
      memcpy(du, w1, sizeof(double) * n);
    } else {
      orc_refdg_update(p, du, x, y, L);
      for (size_t q = 0; q < n; ++q) w1[q] = du[q] + c[0] * dt * L[q];
      orc_refdg_limiter(p, w1);
      orc_refdg_update(p, w1, x, y, L);
      for (size_t q = 0; q < n; ++q) w2[q] = c[1] * du[q] + c[2] * w1[q] + c[3] * dt * L[q];
      orc_refdg_limiter(p, w2);
      orc_refdg_update(p, w2, x, y, L);
      for (size_t q = 0; q < n; ++q) w3[q] = c[4] * du[q] + c[5] * w2[q] + c[6] * dt * L[q];
      orc_refdg_limiter(p, w3);
      orc_refdg_update(p, w3, x, y, L);
      for (size_t q = 0; q < n; ++q) w4[q] = c[7] * du[q] + c[8] * w3[q] + c[9] * dt * L[q];
      orc_refdg_limiter(p, w4);
      orc_refdg_update(p, w3, x, y, L);                   /* L(w3) again, :700 / :723 */
      for (size_t q = 0; q < n; ++q)
        w5[q] = c[10] * du[q] + c[11] * w2[q] + c[12] * w3[q] + c[13] * w4[q] + c[14] * dt * L[q];
      orc_refdg_update(p, w4, x, y, L);
      for (size_t q = 0; q < n; ++q) du[q] = w5[q] + c[15] * dt * L[q];
      orc_refdg_limiter(p, du);
    }
    t = t + dt;
    if (dts) dts[iter] = dt;
    ++iter;
  }
  *t_io = t;
  free(L); free(w1); free(w2); free(w3); free(w4); free(w5);
  return iter;
}

/* compute_error :23-89 on nodal values against u_init: per variable the
 * maximum |u - u_init|, the L1 norm sum_cells (sum_ab |u - u_init| w_a w_b)
 * (dx dy) 0.25 and the L2 norm sqrt(sum_cells (sum_ab (u - u_init)^2 w_a w_b)
 * (dx dy) 0.25), with dx = 1/nx and dy = 1/ny (not the box lengths) and the
 * x weights in both directions, as the reference writes them.
 * out = (linf[4], l1[4], l2[4]). */
void orc_refdg_error(int nxi, int nyi, int oi, const double* u, const double* u_init, double out[12])
{
  const size_t nx = (size_t)nxi, ny = (size_t)nyi, o = (size_t)oi;
  double xq[8], wq[8];
  orc_dg_gl_quadrature(oi, xq, wq);
  const double dx = 1. / (double)nxi, dy = 1. / (double)nyi;
  for (int v = 0; v < 4; ++v) {
    double m = 0.0;
    for (size_t k = 0; k < nx * ny * o * o; ++k) m = fmax_f(m, fabs(u[4 * k + v] - u_init[4 * k + v]));
    out[v] = m;
  }
  for (int v = 0; v < 4; ++v) {
    double l1 = 0.0;
    for (size_t i = 0; i < nx; ++i)
      for (size_t j = 0; j < ny; ++j) {
        double acc = 0.0;
        for (size_t a = 0; a < o; ++a)
          for (size_t b = 0; b < o; ++b)
            acc = acc + fabs(DQ(u, v, i, j, a, b) - DQ(u_init, v, i, j, a, b)) * wq[a] * wq[b];
        l1 = l1 + acc * (dx * dy) * 0.25;
      }
    out[4 + v] = l1;
  }
  for (int v = 0; v < 4; ++v) {
    double l2 = 0.0;
    for (size_t i = 0; i < nx; ++i)
      for (size_t j = 0; j < ny; ++j) {
        double acc = 0.0;
        for (size_t a = 0; a < o; ++a)
          for (size_t b = 0; b < o; ++b) {
            const double d = DQ(u, v, i, j, a, b) - DQ(u_init, v, i, j, a, b);
            acc = acc + d * d * wq[a] * wq[b];
          }
        l2 = l2 + acc * (dx * dy) * 0.25;
      }
    out[8 + v] = sqrt(l2);
  }
}
