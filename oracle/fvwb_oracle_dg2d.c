/*
 * fvwb_oracle_dg2d.c — CPU FP64 oracle, 2D modal DG (TEST INFRASTRUCTURE).
 *
 * Restatements of the reference's DG building blocks, pinned against the
 * flang build of the reference: legendre / legendre_prime / gl_quadrature
 * (2d/legendre.f90:1-108) and the tensor-basis projections
 * get_modes_from_nodes / get_nodes_from_modes (2d/benchmark_2d_dg.f90:497-592).
 *
 * The configured scheme: modal DG of total degree k <= 2 in deviation form
 * (the template of dg_with_source.f90:1749-2031 carried to 2D), Gauss-Legendre
 * volume and face quadrature with the reference's assembly
 * (2d/benchmark_2d_dg.f90:1455-1462), the configured FV face flux, walls, the
 * TVB limiter of 2d/limiters.f90:30-54 on the linear modes, SSP-RK3 or the
 * reference's SSP(5,4) (2d/benchmark_2d_dg.f90:683-710).  Pinned by
 * invariants in tests/ (exact rest, degree 0 == first-order FV, convergence
 * orders, mass identity).  Built with -ffp-contract=off.  Not linked by the
 * product.
 */
#include "fvwb_oracle.h"

#include <math.h>
#include <stdbool.h>
#include <stdlib.h>
#include <string.h>

/* ---- 2d/legendre.f90 ------------------------------------------------------- */

/* legendre(x,n): sqrt(2n+1) P_n(x) with x clamped to [-1,1] (:1-25) */
double orc_dg_legendre(double x, int n)
{
  x = fmin(fmax(x, -1.0), 1.0);
  double x2 = x * x, l = 0.0;
  switch (n) {
  case 0: l = 1.0; break;
  case 1: l = x; break;
  case 2: l = 0.5 * (3 * x2 - 1); break;
  case 3: l = 0.5 * (5.0 * (x2 * x) - 3.0 * x); break;
  case 4: l = 0.125 * (35.0 * (x2 * x2) - 30.0 * x2 + 3.0); break;
  case 5: l = 0.125 * (63.0 * (x2 * x2 * x) - 70.0 * (x2 * x) + 15.0 * x); break;
  case 6: l = 1.0 / 16.0 * (231.0 * (x2 * x2 * x2) - 315.0 * (x2 * x2) + 105.0 * x2 - 5.0); break;
  default: l = 0.0;
  }
  return sqrt(2.0 * (double)n + 1.0) * l;
}

/* legendre_prime(x,n) (:27-50) */
double orc_dg_legendre_prime(double x, int n)
{
  x = fmin(fmax(x, -1.0), 1.0);
  double x2 = x * x, l = 0.0;
  switch (n) {
  case 0: l = 0.0; break;
  case 1: l = 1.0; break;
  case 2: l = 3.0 * x; break;
  case 3: l = 0.5 * (15.0 * x2 - 3.0); break;
  case 4: l = 0.125 * (140.0 * (x2 * x) - 60.0 * x); break;
  case 5: l = 0.125 * (315.0 * (x2 * x2) - 210.0 * x2 + 15.0); break;
  case 6: l = 1.0 / 16.0 * (1386.0 * (x2 * x2 * x) - 1260.0 * (x2 * x) + 210.0 * x); break;
  default: l = 0.0;
  }
  return sqrt(2.0 * (double)n + 1.0) * l;
}

/* gl_quadrature(x_quad,w_quad,n) (:77-108): Newton from the classical
 * initial guess (its coefficient evaluated in default REAL), 500 steps,
 * nodes sign-flipped, weights 2(2n+1)/((1-x^2) L'(x)^2), then mirrored. */
void orc_dg_gl_quadrature(int n, double* xq, double* wq)
{
  const double dpi = acos(-1.0);
  for (int i = 1; i <= n; ++i) {
    float nf = (float)n;
    float c = (1.0f - 0.125f / nf / nf) + 0.125f / nf / nf / nf;
    double xx = (double)c * cos(dpi * (4.0 * (double)i - 1.0) / (4.0 * (double)n + 2.0));
    for (int it = 0; it < 500; ++it) xx = xx - orc_dg_legendre(xx, n) / orc_dg_legendre_prime(xx, n);
    xq[i - 1] = -xx;
    double lp = orc_dg_legendre_prime(xq[i - 1], n);
    wq[i - 1] = 2 * (2.0 * (double)n + 1.0) / (1.0 - xq[i - 1] * xq[i - 1]) / (lp * lp);
  }
  for (int i = n / 2 + 1; i <= n; ++i) {
    xq[i - 1] = -xq[n - i];
    wq[i - 1] = wq[n - i];
  }
}

/* ---- 2d/benchmark_2d_dg.f90:497-592, tensor (Q) basis ------------------------ */
/* layout AoS(4, nx, ny, order, order): element (v,i,j,a,b) */
static inline size_t q5(int v, int i, int j, int a, int b, int nx, int ny, int o)
{
  return (size_t)v + 4u * ((size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * ((size_t)a + (size_t)o * (size_t)b)));
}
#define Q5(arr, v, i, j, a, b, nx, ny, o) ((arr)[q5((v), (i), (j), (a), (b), (nx), (ny), (o))])

void orc_ref_dg2d_modes_from_nodes(int nx, int ny, int order, const double* nodes, double* modes)
{
  double xq[8], wq[8];
  orc_dg_gl_quadrature(order, xq, wq);
  memset(modes, 0, sizeof(double) * 4 * (size_t)nx * ny * order * order);
  for (int ic = 0; ic < nx; ++ic)
    for (int jc = 0; jc < ny; ++jc)
      for (int i = 0; i < order; ++i)
        for (int j = 0; j < order; ++j)
          for (int a = 0; a < order; ++a)
            for (int b = 0; b < order; ++b)
              for (int v = 0; v < 4; ++v)
                Q5(modes, v, ic, jc, i, j, nx, ny, order) =
                    Q5(modes, v, ic, jc, i, j, nx, ny, order) +
                    0.25 * Q5(nodes, v, ic, jc, a, b, nx, ny, order) * orc_dg_legendre(xq[a], i) *
                        orc_dg_legendre(xq[b], j) * wq[a] * wq[b];
}

void orc_ref_dg2d_nodes_from_modes(int nx, int ny, int order, const double* modes, double* nodes)
{
  double xq[8], wq[8];
  orc_dg_gl_quadrature(order, xq, wq);
  memset(nodes, 0, sizeof(double) * 4 * (size_t)nx * ny * order * order);
  for (int v = 0; v < 4; ++v)
    for (int ic = 0; ic < nx; ++ic)
      for (int jc = 0; jc < ny; ++jc)
        for (int i = 0; i < order; ++i)
          for (int j = 0; j < order; ++j)
            for (int a = 0; a < order; ++a)
              for (int b = 0; b < order; ++b)
                Q5(nodes, v, ic, jc, i, j, nx, ny, order) =
                    Q5(nodes, v, ic, jc, i, j, nx, ny, order) +
                    Q5(modes, v, ic, jc, a, b, nx, ny, order) * orc_dg_legendre(xq[i], a) *
                        orc_dg_legendre(xq[j], b);
}

/* ---- configured scheme ------------------------------------------------------ */

typedef struct {
  int k, nm, nq;
  int mi[6], mj[6];
  double xq[3], wq[3];
  double phi[3][3];   /* phi[n][a] = L_n(xq[a]) */
  double dphi[3][3];  /* L_n'(xq[a]) */
  double phim[3], phip[3];  /* L_n(-1), L_n(+1) */
} dg_basis;

static const int MODE_I[6] = {0, 1, 0, 2, 1, 0};
static const int MODE_J[6] = {0, 0, 1, 0, 1, 2};

int orc_dg2d_nmodes(int degree) { return (degree + 1) * (degree + 2) / 2; }

static void make_basis(int k, dg_basis* B)
{
  B->k = k;
  B->nm = orc_dg2d_nmodes(k);
  B->nq = k + 1;
  for (int m = 0; m < B->nm; ++m) { B->mi[m] = MODE_I[m]; B->mj[m] = MODE_J[m]; }
  orc_dg_gl_quadrature(B->nq, B->xq, B->wq);
  for (int n = 0; n <= k; ++n) {
    for (int a = 0; a < B->nq; ++a) {
      B->phi[n][a] = orc_dg_legendre(B->xq[a], n);
      B->dphi[n][a] = orc_dg_legendre_prime(B->xq[a], n);
    }
    B->phim[n] = orc_dg_legendre(-1.0, n);
    B->phip[n] = orc_dg_legendre(1.0, n);
  }
}

/* host layout of modes: AoS(4, nx, ny, nm), element (v,i,j,m) */
#define DM(arr, v, i, j, m, nx, ny) \
  ((arr)[(size_t)(v) + 4u * ((size_t)(i) + (size_t)(nx) * ((size_t)(j) + (size_t)(ny) * (size_t)(m)))])

/* the FV parameter block that carries the face flux settings */
static fvwb_params fv_of(const fvwb_dg_params* p)
{
  fvwb_params q;
  memset(&q, 0, sizeof q);
  q.nvar = 4; q.nx = p->nx; q.ny = p->ny;
  q.bc = FVWB_BC_REFLECT; q.recon = FVWB_RECON_FIRST_ORDER; q.flux = p->flux;
  q.integrator = FVWB_SSPRK3; q.precision = FVWB_FP64;
  q.gamma = p->gamma; q.boxlen_x = p->boxlen_x; q.boxlen_y = p->boxlen_y; q.cfl = p->cfl;
  q.rho0 = p->rho0; q.p0 = p->p0; q.g = p->g;
  return q;
}

/* conservative equilibrium at rest at (x,y) */
static void ueq_at(const fvwb_dg_params* p, double x, double y, double u[4])
{
  const double kk = p->rho0 * p->g / p->p0;
  const double e = exp(-(kk * (x + y)));
  u[0] = p->rho0 * e;
  u[1] = 0.0;
  u[2] = 0.0;
  u[3] = p->p0 * e / (p->gamma - 1.0);
}

/* deviation of cell (i,j) at (xi, eta) given as basis values per direction */
static void eval_dev(const double* d, int i, int j, int nx, int ny, const dg_basis* B, const double* lx,
                     const double* ly, double out[4])
{
  for (int v = 0; v < 4; ++v) {
    double s = 0.0;
    for (int m = 0; m < B->nm; ++m) s += DM(d, v, i, j, m, nx, ny) * (lx[B->mi[m]] * ly[B->mj[m]]);
    out[v] = s;
  }
}

void orc_dg2d_residual(const fvwb_dg_params* p, const double* d, double* L)
{
  const int nx = p->nx, ny = p->ny;
  dg_basis B;
  make_basis(p->degree, &B);
  const int nm = B.nm, nq = B.nq;
  const double dx = p->boxlen_x / nx, dy = p->boxlen_y / ny;
  const double odx = 1.0 / dx, ody = 1.0 / dy;
  const fvwb_params fv = fv_of(p);
  memset(L, 0, sizeof(double) * 4 * (size_t)nx * ny * nm);
  double lx[3], ly[3];

  /* volume and source integrals */
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const double xc = (i + 0.5) * dx, yc = (j + 0.5) * dy;
      for (int b = 0; b < nq; ++b)
        for (int a = 0; a < nq; ++a) {
          for (int n = 0; n <= B.k; ++n) { lx[n] = B.phi[n][a]; ly[n] = B.phi[n][b]; }
          double dv[4], ue[4], uu[4], fx[4], fy[4], fxe[4], fye[4];
          eval_dev(d, i, j, nx, ny, &B, lx, ly, dv);
          ueq_at(p, xc + 0.5 * dx * B.xq[a], yc + 0.5 * dy * B.xq[b], ue);
          for (int v = 0; v < 4; ++v) uu[v] = ue[v] + dv[v];
          orc_ref_flux(p->gamma, uu, fx, fy);
          orc_ref_flux(p->gamma, ue, fxe, fye);
          const double s[4] = {0.0, -p->g * dv[0], -p->g * dv[0], -p->g * (dv[1] + dv[2])};
          const double w = B.wq[a] * B.wq[b];
          for (int m = 0; m < nm; ++m) {
            const int mi = B.mi[m], mj = B.mj[m];
            const double gx = 0.5 * w * B.dphi[mi][a] * B.phi[mj][b] * odx;
            const double gy = 0.5 * w * B.phi[mi][a] * B.dphi[mj][b] * ody;
            const double gs = 0.25 * w * B.phi[mi][a] * B.phi[mj][b];
            for (int v = 0; v < 4; ++v)
              DM(L, v, i, j, m, nx, ny) += gx * (fx[v] - fxe[v]) + gy * (fy[v] - fye[v]) + gs * s[v];
          }
        }
    }

  /* x faces: face fi between cells fi-1 and fi */
  for (int j = 0; j < ny; ++j)
    for (int fi = 0; fi <= nx; ++fi)
      for (int b = 0; b < nq; ++b) {
        double dl[4], dr[4], ue[4], f[4];
        for (int n = 0; n <= B.k; ++n) ly[n] = B.phi[n][b];
        if (fi > 0) eval_dev(d, fi - 1, j, nx, ny, &B, B.phip, ly, dl);
        if (fi < nx) eval_dev(d, fi, j, nx, ny, &B, B.phim, ly, dr);
        if (fi == 0) { memcpy(dl, dr, sizeof dl); dl[1] = -dl[1]; }
        if (fi == nx) { memcpy(dr, dl, sizeof dr); dr[1] = -dr[1]; }
        ueq_at(p, fi * dx, (j + 0.5) * dy + 0.5 * dy * B.xq[b], ue);
        orc_fv2d_face_flux(&fv, 0, ue, dl, dr, f);
        const double w = 0.5 * B.wq[b] * odx;
        for (int m = 0; m < nm; ++m) {
          const double pj = B.phi[B.mj[m]][b];
          for (int v = 0; v < 4; ++v) {
            if (fi > 0) DM(L, v, fi - 1, j, m, nx, ny) -= w * f[v] * B.phip[B.mi[m]] * pj;
            if (fi < nx) DM(L, v, fi, j, m, nx, ny) += w * f[v] * B.phim[B.mi[m]] * pj;
          }
        }
      }

  /* y faces: face fj between cells fj-1 and fj */
  for (int fj = 0; fj <= ny; ++fj)
    for (int i = 0; i < nx; ++i)
      for (int a = 0; a < nq; ++a) {
        double dl[4], dr[4], ue[4], f[4];
        for (int n = 0; n <= B.k; ++n) lx[n] = B.phi[n][a];
        if (fj > 0) eval_dev(d, i, fj - 1, nx, ny, &B, lx, B.phip, dl);
        if (fj < ny) eval_dev(d, i, fj, nx, ny, &B, lx, B.phim, dr);
        if (fj == 0) { memcpy(dl, dr, sizeof dl); dl[2] = -dl[2]; }
        if (fj == ny) { memcpy(dr, dl, sizeof dr); dr[2] = -dr[2]; }
        ueq_at(p, (i + 0.5) * dx + 0.5 * dx * B.xq[a], fj * dy, ue);
        orc_fv2d_face_flux(&fv, 1, ue, dl, dr, f);
        const double w = 0.5 * B.wq[a] * ody;
        for (int m = 0; m < nm; ++m) {
          const double pi_ = B.phi[B.mi[m]][a];
          for (int v = 0; v < 4; ++v) {
            if (fj > 0) DM(L, v, i, fj - 1, m, nx, ny) -= w * f[v] * pi_ * B.phip[B.mj[m]];
            if (fj < ny) DM(L, v, i, fj, m, nx, ny) += w * f[v] * pi_ * B.phim[B.mj[m]];
          }
        }
      }
}

/* TVB limiter on the linear modes of each conservative deviation; the means
 * are never changed, so cells can be limited in any order. */
/* The positivity limiter of the configured scheme: compute_positivity
 * (2d/limiters.f90:478-654) applied to the full state's modes F = E + d, E the
 * (k+1)x(k+1) Gauss projection of U_eq on the cell's P_k basis.  The set is the
 * reference's compute_set for k+1 Gauss points, (x_gll, y_gl) and (x_gl, y_gll)
 * with the k+1 Gauss-Lobatto points (-1, 0, 1 for k = 2; -1, 1 for k = 1), plus
 * the volume Gauss points, where the scheme evaluates the state too.  The
 * thresholds are the cell's: 1e-5 of its equilibrium mean density and pressure
 * (at least the reference's 1e-10), which an FP32 deviation resolves.  A
 * variable whose modes the limiter leaves alone keeps its deviation bit for bit. */
static void positivity_cell(const fvwb_dg_params* p, const dg_basis* B, int i, int j, double* d)
{
  const int nx = p->nx, ny = p->ny, nm = B->nm, nq = B->nq;
  const double dx = p->boxlen_x / nx, dy = p->boxlen_y / ny;
  double E[4][6], F[4][6];
  for (int v = 0; v < 4; ++v)
    for (int m = 0; m < nm; ++m) E[v][m] = 0.0;
  for (int b = 0; b < nq; ++b)
    for (int a = 0; a < nq; ++a) {
      double ue[4];
      ueq_at(p, (i + 0.5) * dx + 0.5 * dx * B->xq[a], (j + 0.5) * dy + 0.5 * dy * B->xq[b], ue);
      for (int m = 0; m < nm; ++m) {
        const double w = 0.25 * B->wq[a] * B->wq[b] * B->phi[B->mi[m]][a] * B->phi[B->mj[m]][b];
        for (int v = 0; v < 4; ++v) E[v][m] += w * ue[v];
      }
    }
  for (int v = 0; v < 4; ++v)
    for (int m = 0; m < nm; ++m) F[v][m] = E[v][m] + DM(d, v, i, j, m, nx, ny);
  const double eps_r = fmax((double)1e-10f, 1e-5 * E[0][0]);
  const double eps_p = fmax((double)1e-10f, 1e-5 * (p->gamma - 1.0) * E[3][0]);
  /* the set: basis values per direction at Gauss-Lobatto (g) and Gauss (q) points */
  double lg[3][3], lq[3][3], xg[3];
  if (nq == 2) { xg[0] = -1.0; xg[1] = 1.0; } else { xg[0] = -1.0; xg[1] = 0.0; xg[2] = 1.0; }
  for (int c = 0; c < nq; ++c)
    for (int n = 0; n <= B->k; ++n) { lg[c][n] = orc_dg_legendre(xg[c], n); lq[c][n] = B->phi[n][c]; }
  const int np = 3 * nq * nq;
  double pts[27][4];
  for (int pass = 0; pass < 2; ++pass) {
    for (int q = 0; q < nq; ++q)
      for (int c = 0; c < nq; ++c)
        for (int side = 0; side < 3; ++side) {
          const double* lx = side == 0 ? lg[c] : lq[q];
          const double* ly = side == 0 ? lq[q] : (side == 1 ? lg[c] : lq[c]);
          double* o = pts[(q * nq + c) + side * nq * nq];
          for (int v = 0; v < 4; ++v) {
            double s = 0.0;
            for (int m = 0; m < nm; ++m) s += F[v][m] * (lx[B->mi[m]] * ly[B->mj[m]]);
            o[v] = s;
          }
        }
    if (pass == 0) {
      double rmin = pts[0][0];
      for (int k = 1; k < np; ++k) rmin = fmin(rmin, pts[k][0]);
      const double theta = fmin(fabs((F[0][0] - eps_r) / (F[0][0] - rmin)), 1.0);
      if (theta < 1.0) {
        for (int m = 1; m < nm; ++m) {
          F[0][m] = theta * F[0][m];
          DM(d, 0, i, j, m, nx, ny) = F[0][m] - E[0][m];
        }
      }
    } else {
      const double avg[4] = {F[0][0], F[1][0], F[2][0], F[3][0]};
      double tmin = 1.0;
      for (int k = 0; k < np; ++k) {
        const double r = fmax(pts[k][0], (double)10e-10f);
        const double vx = pts[k][1] / r, vy = pts[k][2] / r;
        const double pr = (p->gamma - 1.0) * (pts[k][3] - 0.5 * r * (vx * vx + vy * vy));
        const double t = pr > eps_p ? 1.0 : orc_ref_dg2d_solve_for_t(p->gamma, eps_p, pts[k], avg);
        tmin = fmin(tmin, t);
      }
      if (tmin < 1.0)
        for (int v = 0; v < 4; ++v)
          for (int m = 1; m < nm; ++m) DM(d, v, i, j, m, nx, ny) = tmin * F[v][m] - E[v][m];
    }
  }
}

/* Eigenvectors of the 2D Euler flux Jacobian along x (dir 0) or y (dir 1) at
 * the conservative state u: columns of R are the right eigenvectors for the
 * speeds vn - c, vn, vn (shear), vn + c; L = R^-1 written in closed form. */
void orc_euler_eigenvectors(double gamma, int dir, const double u[4], double R[16], double L[16])
{
  const double rho = u[0], vx = u[1] / rho, vy = u[2] / rho;
  const double q2 = vx * vx + vy * vy;
  const double p = (gamma - 1.0) * (u[3] - 0.5 * rho * q2);
  const double c = sqrt(gamma * p / rho), H = (u[3] + p) / rho;
  const double vn = dir == 0 ? vx : vy, vt = dir == 0 ? vy : vx;
  const int n = dir == 0 ? 1 : 2, t = dir == 0 ? 2 : 1;   /* normal and tangential momentum rows */
  const double b1 = (gamma - 1.0) / (c * c), b2 = 0.5 * b1 * q2;
  double r[4][4] = {{0}}, l[4][4] = {{0}};
  /* right eigenvectors (rows = conserved variable, columns = wave) */
  r[0][0] = 1.0;     r[0][1] = 1.0;      r[0][2] = 0.0; r[0][3] = 1.0;
  r[n][0] = vn - c;  r[n][1] = vn;       r[n][2] = 0.0; r[n][3] = vn + c;
  r[t][0] = vt;      r[t][1] = vt;       r[t][2] = 1.0; r[t][3] = vt;
  r[3][0] = H - vn * c; r[3][1] = 0.5 * q2; r[3][2] = vt; r[3][3] = H + vn * c;
  /* left eigenvectors (rows = wave) */
  l[0][0] = 0.5 * (b2 + vn / c); l[0][n] = -0.5 * (b1 * vn + 1.0 / c); l[0][t] = -0.5 * b1 * vt; l[0][3] = 0.5 * b1;
  l[1][0] = 1.0 - b2;            l[1][n] = b1 * vn;                  l[1][t] = b1 * vt;         l[1][3] = -b1;
  l[2][0] = -vt;                 l[2][n] = 0.0;                      l[2][t] = 1.0;             l[2][3] = 0.0;
  l[3][0] = 0.5 * (b2 - vn / c); l[3][n] = -0.5 * (b1 * vn - 1.0 / c); l[3][t] = -0.5 * b1 * vt; l[3][3] = 0.5 * b1;
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 4; ++k) { R[4 * i + k] = r[i][k]; L[4 * i + k] = l[i][k]; }
}

static void mat4(const double* M, const double* x, double* y)
{
  for (int i = 0; i < 4; ++i) y[i] = M[4 * i] * x[0] + M[4 * i + 1] * x[1] + M[4 * i + 2] * x[2] + M[4 * i + 3] * x[3];
}

/* characteristic TVB limiting of the linear modes (limiter 4, 5) */
static void characteristic_limit(const fvwb_dg_params* p, double* d)
{
  const int nx = p->nx, ny = p->ny, nm = orc_dg2d_nmodes(p->degree);
  const double dx = p->boxlen_x / nx, dy = p->boxlen_y / ny;
  const double s3 = sqrt(3.0);
  dg_basis B;
  make_basis(p->degree, &B);
  double* out = malloc(sizeof(double) * 4 * (size_t)nx * ny * nm);
  memcpy(out, d, sizeof(double) * 4 * (size_t)nx * ny * nm);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      /* the full-state mean: the equilibrium's Gauss-projected mean plus the deviation's */
      double ubar[4] = {0.0, 0.0, 0.0, 0.0};
      for (int b = 0; b < B.nq; ++b)
        for (int a = 0; a < B.nq; ++a) {
          double ue[4];
          ueq_at(p, (i + 0.5) * dx + 0.5 * dx * B.xq[a], (j + 0.5) * dy + 0.5 * dy * B.xq[b], ue);
          for (int v = 0; v < 4; ++v) ubar[v] += 0.25 * B.wq[a] * B.wq[b] * ue[v];
        }
      for (int v = 0; v < 4; ++v) ubar[v] += DM(d, v, i, j, 0, nx, ny);
      bool changed = false;
      double lin[2][4];
      for (int dir = 0; dir < 2; ++dir) {
        double R[16], L[16], a[4], dp[4], dm[4], ca[4], cp[4], cm[4];
        orc_euler_eigenvectors(p->gamma, dir, ubar, R, L);
        for (int v = 0; v < 4; ++v) {
          const double u0 = DM(d, v, i, j, 0, nx, ny);
          double up, um;
          if (dir == 0) {
            up = i < nx - 1 ? DM(d, v, i + 1, j, 0, nx, ny) : (v == 1 ? -u0 : u0);
            um = i > 0 ? DM(d, v, i - 1, j, 0, nx, ny) : (v == 1 ? -u0 : u0);
          } else {
            up = j < ny - 1 ? DM(d, v, i, j + 1, 0, nx, ny) : (v == 2 ? -u0 : u0);
            um = j > 0 ? DM(d, v, i, j - 1, 0, nx, ny) : (v == 2 ? -u0 : u0);
          }
          a[v] = s3 * DM(d, v, i, j, 1 + dir, nx, ny);
          dp[v] = up - u0;
          dm[v] = u0 - um;
        }
        mat4(L, a, ca);
        mat4(L, dp, cp);
        mat4(L, dm, cm);
        bool lim = false;
        double cl[4];
        for (int k = 0; k < 4; ++k) {
          cl[k] = orc_generalized_minmod(ca[k], cp[k], cm[k], p->tvb_m, dir == 0 ? dx : dy);
          lim |= cl[k] != ca[k];
        }
        if (lim) {
          double a2[4];
          mat4(R, cl, a2);
          for (int v = 0; v < 4; ++v) lin[dir][v] = a2[v] / s3;
          changed = true;
        } else {
          for (int v = 0; v < 4; ++v) lin[dir][v] = DM(d, v, i, j, 1 + dir, nx, ny);
        }
      }
      if (changed)
        for (int v = 0; v < 4; ++v) {
          DM(out, v, i, j, 1, nx, ny) = lin[0][v];
          DM(out, v, i, j, 2, nx, ny) = lin[1][v];
          for (int m = 3; m < nm; ++m) DM(out, v, i, j, m, nx, ny) = 0.0;
        }
    }
  memcpy(d, out, sizeof(double) * 4 * (size_t)nx * ny * nm);
  free(out);
}

void orc_dg2d_limit(const fvwb_dg_params* p, double* d)
{
  if (p->degree < 1) return;
  if (p->limiter == FVWB_DG_LIMITER_CHARACTERISTIC || p->limiter == FVWB_DG_LIMITER_CHARACTERISTIC_POSITIVITY) {
    characteristic_limit(p, d);
    if (p->limiter == FVWB_DG_LIMITER_CHARACTERISTIC_POSITIVITY) {
      dg_basis B;
      make_basis(p->degree, &B);
      for (int j = 0; j < p->ny; ++j)
        for (int i = 0; i < p->nx; ++i) positivity_cell(p, &B, i, j, d);
    }
    return;
  }
  if (p->limiter == FVWB_DG_LIMITER_POSITIVITY || p->limiter == FVWB_DG_LIMITER_TVB_POSITIVITY) {
    if (p->limiter == FVWB_DG_LIMITER_TVB_POSITIVITY) {
      fvwb_dg_params q = *p;
      q.limiter = FVWB_DG_LIMITER_TVB;
      orc_dg2d_limit(&q, d);
    }
    dg_basis B;
    make_basis(p->degree, &B);
    for (int j = 0; j < p->ny; ++j)
      for (int i = 0; i < p->nx; ++i) positivity_cell(p, &B, i, j, d);
    return;
  }
  if (p->limiter != FVWB_DG_LIMITER_TVB) return;
  const int nx = p->nx, ny = p->ny, nm = orc_dg2d_nmodes(p->degree);
  const double dx = p->boxlen_x / nx, dy = p->boxlen_y / ny;
  const double s3 = sqrt(3.0);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i)
      for (int v = 0; v < 4; ++v) {
        const double u0 = DM(d, v, i, j, 0, nx, ny);
        /* neighbour means; walls mirror the cell, normal momentum negated */
        const double uw = i > 0 ? DM(d, v, i - 1, j, 0, nx, ny) : (v == 1 ? -u0 : u0);
        const double ue = i < nx - 1 ? DM(d, v, i + 1, j, 0, nx, ny) : (v == 1 ? -u0 : u0);
        const double us = j > 0 ? DM(d, v, i, j - 1, 0, nx, ny) : (v == 2 ? -u0 : u0);
        const double un = j < ny - 1 ? DM(d, v, i, j + 1, 0, nx, ny) : (v == 2 ? -u0 : u0);
        const double a = s3 * DM(d, v, i, j, 1, nx, ny);
        const double b = s3 * DM(d, v, i, j, 2, nx, ny);
        const double at = orc_generalized_minmod(a, ue - u0, u0 - uw, p->tvb_m, dx);
        const double bt = orc_generalized_minmod(b, un - u0, u0 - us, p->tvb_m, dy);
        if (at != a || bt != b) {
          DM(d, v, i, j, 1, nx, ny) = at / s3;
          DM(d, v, i, j, 2, nx, ny) = bt / s3;
          for (int m = 3; m < nm; ++m) DM(d, v, i, j, m, nx, ny) = 0.0;
        }
      }
}

/* ---- the reference's positivity limiter (2d/limiters.f90:312-654), as written ---- */
/* 2d/legendre.f90:111-170 gll_quadrature: n = 2 gives the points (-1, 0), n = 3
 * (-1, 0, 1), as written; only the points are used by compute_set */
void orc_ref_dg2d_gll_points(int n, double* x)
{
  if (n == 2) { x[0] = -1.; x[1] = 0.; }
  else if (n == 3) { x[0] = -1.; x[1] = 0.; x[2] = 1.; }
  else for (int i = 0; i < n; ++i) x[i] = NAN;
}

/* 2d/limiters.f90:312-362 solve_for_t: the root in [0,1] of the quadratic whose
 * zero puts the pressure of (1-t) u_avg + t u at eps, written with the
 * reference's operation order (a, b, c, then b/a, c/a); no root in range gives 0 */
double orc_ref_dg2d_solve_for_t(double gamma, double eps, const double u[4], const double ua[4])
{
  const double pa = ua[0], mxa = ua[1], mya = ua[2], ea = ua[3];
  const double pj = u[0], mxj = u[1], myj = u[2], ej = u[3];
  double a = 2.0 * (pj - pa) * (ej - ea) - (mxj - mxa) * (mxj - mxa) - (myj - mya) * (myj - mya);
  double b = 2.0 * (pj - pa) * (ea - eps / (gamma - 1)) + 2.0 * pa * (ej - ea) -
             2.0 * (mxa * (mxj - mxa) + mya * (myj - mya));
  double c = 2.0 * pa * ea - (mxa * mxa + mya * mya) - 2.0 * eps * pa / (gamma - 1.0);
  b = b / a;
  c = c / a;
  const double D = sqrt(fabs(b * b - 4 * c));
  const double t1 = 0.5 * (-b - D), t2 = 0.5 * (-b + D);
  double t;
  if (t1 > -eps && t1 < 1.0 + eps) t = t1;
  else if (t2 > -eps && t2 < 1.0 + eps) t = t2;
  else t = 0.0;
  t = fmin(1.0, t);
  t = fmax(0.0, t);
  return t;
}

/* 2d/limiters.f90:438-475 compute_set for one cell's modes m(v,a,b) (order o,
 * gll points): the values at (x_gll(c), y_gl(q)) then at (x_gl(q), x_gll(c)),
 * point index q*gll + c and then + o*gll */
static void ref_positivity_set(const double* m, int o, int gll, double* pts)
{
  double xq[8], wq[8], xg[8];
  orc_dg_gl_quadrature(o, xq, wq);
  orc_ref_dg2d_gll_points(gll, xg);
  for (int q = 0; q < o; ++q)
    for (int c = 0; c < gll; ++c) {
      double ul[4] = {0.0, 0.0, 0.0, 0.0}, ur[4] = {0.0, 0.0, 0.0, 0.0};
      for (int i = 0; i < o; ++i)
        for (int j = 0; j < o; ++j)
          for (int v = 0; v < 4; ++v) {
            const double mv = m[v + 4 * (i + o * j)];
            ul[v] = ul[v] + mv * orc_dg_legendre(xg[c], i) * orc_dg_legendre(xq[q], j);
            ur[v] = ur[v] + mv * orc_dg_legendre(xq[q], i) * orc_dg_legendre(xg[c], j);
          }
      for (int v = 0; v < 4; ++v) {
        pts[v + 4 * (q * gll + c)] = ul[v];
        pts[v + 4 * (q * gll + c + o * gll)] = ur[v];
      }
    }
}

/* 2d/limiters.f90:478-654 compute_positivity on the tensor modes
 * u(1:4,1:nx,1:ny,1:o,1:o) of the reference's own DG, in place: density first,
 * theta = min(|(rho_avg - eps)/(rho_avg - rho_min)|, 1) over the set, scaling
 * the non-mean density modes; then, on the density-limited modes, the smallest
 * solve_for_t over the set points whose pressure (density floored at 1e-9) is
 * not above eps, scaling every non-mean mode.  gll = (2(o-1)+3)/2. */
void orc_ref_dg2d_positivity(int nx, int ny, int o, double gamma, double eps, double* u)
{
  if (o == 1) return;
  const int gll = (2 * (o - 1) + 3) / 2, np = 2 * o * gll;
  double m[4 * 64], pts[4 * 64];
  for (int pass = 0; pass < 2; ++pass)
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) {
        for (int a = 0; a < o; ++a)
          for (int b = 0; b < o; ++b)
            for (int v = 0; v < 4; ++v) m[v + 4 * (a + o * b)] = Q5(u, v, i, j, a, b, nx, ny, o);
        ref_positivity_set(m, o, gll, pts);
        const double avg[4] = {m[0], m[1], m[2], m[3]};
        if (pass == 0) {
          double pmin = pts[0];
          for (int k = 1; k < np; ++k) pmin = fmin(pmin, pts[4 * k]);
          const double theta = fmin(fabs((avg[0] - eps) / (avg[0] - pmin)), 1.0);
          for (int a = 0; a < o; ++a)
            for (int b = 0; b < o; ++b)
              if (a != 0 || b != 0) Q5(u, 0, i, j, a, b, nx, ny, o) = theta * m[4 * (a + o * b)];
        } else {
          double tmin = 1.;
          for (int k = 0; k < np; ++k) {
            const double* pk = &pts[4 * k];
            const double r = fmax(pk[0], (double)10e-10f);
            const double vx = pk[1] / r, vy = pk[2] / r;
            const double pr = (gamma - 1.0) * (pk[3] - 0.5 * r * (vx * vx + vy * vy));
            const double t = pr > eps ? 1. : orc_ref_dg2d_solve_for_t(gamma, eps, pk, avg);
            if (tmin >= t) tmin = t;
          }
          for (int a = 0; a < o; ++a)
            for (int b = 0; b < o; ++b)
              if (a != 0 || b != 0)
                for (int v = 0; v < 4; ++v) Q5(u, v, i, j, a, b, nx, ny, o) = tmin * m[v + 4 * (a + o * b)];
        }
      }
}

/* max over cells and volume points of (|u|+c)/dx + (|v|+c)/dy of U_eq + dU */
double orc_dg2d_max_rate(const fvwb_dg_params* p, const double* d)
{
  const int nx = p->nx, ny = p->ny;
  dg_basis B;
  make_basis(p->degree, &B);
  const double dx = p->boxlen_x / nx, dy = p->boxlen_y / ny;
  double r = 0.0, lx[3], ly[3];
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i)
      for (int b = 0; b < B.nq; ++b)
        for (int a = 0; a < B.nq; ++a) {
          for (int n = 0; n <= B.k; ++n) { lx[n] = B.phi[n][a]; ly[n] = B.phi[n][b]; }
          double dv[4], ue[4], u[4], w[4];
          eval_dev(d, i, j, nx, ny, &B, lx, ly, dv);
          ueq_at(p, (i + 0.5) * dx + 0.5 * dx * B.xq[a], (j + 0.5) * dy + 0.5 * dy * B.xq[b], ue);
          for (int v = 0; v < 4; ++v) u[v] = ue[v] + dv[v];
          orc_ref_primitive(p->gamma, u, w);
          const double c = sqrt(p->gamma * fmax(w[3], 1e-10) / fmax(w[0], 1e-10));
          r = fmax(r, (fabs(w[1]) + c) / dx + (fabs(w[2]) + c) / dy);
        }
  return r;
}

/* SSP-RK3 with the limiter after every stage; the state is taken as limited
 * on entry.  dt = cfl/((2k+1) max rate), the last step clipped to t_end. */
/* The five-stage fourth-order SSP Runge-Kutta of 2d/benchmark_2d_dg.f90:683-710
 * (solver 'RK4'), in its Shu-Osher form.  The reference writes the coefficients
 * as default-real literals, so they enter the FP64 arithmetic rounded to FP32;
 * they are kept that way here:
 *   w1 = u + c10 dt L(u)
 *   w2 = a20 u + a21 w1 + c21 dt L(w1)
 *   w3 = a30 u + a32 w2 + c32 dt L(w2)
 *   w4 = a40 u + a43 w3 + c43 dt L(w3)
 *   u' = a50 u + a52 w2 + a53 w3 + a54 w4 + c53 dt L(w3) + c54 dt L(w4)
 * in the order c10, a20, a21, c21, a30, a32, c32, a40, a43, c43, a50, a52, a53,
 * a54, c53, c54. */
void orc_dg2d_ssp54_coeffs(double c[16])
{
  static const float lit[16] = {
      0.391752226571890f,                                              /* :686 */
      0.444370493651235f, 0.555629506348765f, 0.368410593050371f,      /* :690 */
      0.620101851488403f, 0.379898148511597f, 0.251891774271694f,      /* :693 */
      0.178079954393132f, 0.821920045606868f, 0.544974750228521f,      /* :696 */
      0.00683325884039f, 0.51723167208978f, 0.12759831133288f, 0.34833675773694f,
      0.08460416338212f,                                               /* :703-704 */
      0.22600748319395f};                                              /* :707 */
  for (int k = 0; k < 16; ++k) c[k] = (double)lit[k];
}

void orc_dg2d_advance(const fvwb_dg_params* p, double* d, double t_end, int max_steps, double* t_out,
                      int* steps_out, double* dt_last)
{
  const size_t n = 4u * (size_t)p->nx * p->ny * orc_dg2d_nmodes(p->degree);
  const int ssp54 = p->integrator == FVWB_DG_SSPRK54;
  double* L = malloc(sizeof(double) * n);
  double* d1 = malloc(sizeof(double) * n);
  double* d2 = malloc(sizeof(double) * n);
  double* d3 = ssp54 ? malloc(sizeof(double) * n) : NULL;
  double* d4 = ssp54 ? malloc(sizeof(double) * n) : NULL;
  double* d5 = ssp54 ? malloc(sizeof(double) * n) : NULL;
  double c[16];
  orc_dg2d_ssp54_coeffs(c);
  double t = 0.0, dt = 0.0;
  int iter = 0;
  while (t < t_end && iter < max_steps) {
    const double rate = orc_dg2d_max_rate(p, d);
    dt = p->cfl / ((2.0 * p->degree + 1.0) * rate);
    if (t + dt > t_end) dt = t_end - t;
    if (!ssp54) {
      orc_dg2d_residual(p, d, L);
      for (size_t q = 0; q < n; ++q) d1[q] = d[q] + dt * L[q];
      orc_dg2d_limit(p, d1);
      orc_dg2d_residual(p, d1, L);
      for (size_t q = 0; q < n; ++q) d2[q] = 0.75 * d[q] + 0.25 * d1[q] + 0.25 * dt * L[q];
      orc_dg2d_limit(p, d2);
      orc_dg2d_residual(p, d2, L);
      for (size_t q = 0; q < n; ++q) d[q] = (1.0 / 3.0) * d[q] + (2.0 / 3.0) * d2[q] + (2.0 / 3.0) * dt * L[q];
      orc_dg2d_limit(p, d);
    } else {
      /* :685-708; L(w3) is evaluated twice there (:695, :700) with the same
       * argument, once here */
      orc_dg2d_residual(p, d, L);
      for (size_t q = 0; q < n; ++q) d1[q] = d[q] + c[0] * dt * L[q];
      orc_dg2d_limit(p, d1);
      orc_dg2d_residual(p, d1, L);
      for (size_t q = 0; q < n; ++q) d2[q] = c[1] * d[q] + c[2] * d1[q] + c[3] * dt * L[q];
      orc_dg2d_limit(p, d2);
      orc_dg2d_residual(p, d2, L);
      for (size_t q = 0; q < n; ++q) d3[q] = c[4] * d[q] + c[5] * d2[q] + c[6] * dt * L[q];
      orc_dg2d_limit(p, d3);
      orc_dg2d_residual(p, d3, L);
      for (size_t q = 0; q < n; ++q) d4[q] = c[7] * d[q] + c[8] * d3[q] + c[9] * dt * L[q];
      orc_dg2d_limit(p, d4);
      /* w5 takes the limited w4 (apply_limiter(w4) precedes it) and L(w3) */
      for (size_t q = 0; q < n; ++q)
        d5[q] = c[10] * d[q] + c[11] * d2[q] + c[12] * d3[q] + c[13] * d4[q] + c[14] * dt * L[q];
      orc_dg2d_residual(p, d4, L);
      for (size_t q = 0; q < n; ++q) d[q] = d5[q] + c[15] * dt * L[q];
      orc_dg2d_limit(p, d);
    }
    t = t + dt;
    iter = iter + 1;
  }
  if (t_out) *t_out = t;
  if (steps_out) *steps_out = iter;
  if (dt_last) *dt_last = dt;
  free(L); free(d1); free(d2); free(d3); free(d4); free(d5);
}

/* values of the deviation at the (k+1)^2 volume points, AoS(4,nx,ny,nq*nq)
 * with point index a + nq*b */
void orc_dg2d_nodal(const fvwb_dg_params* p, const double* d, double* nodal)
{
  const int nx = p->nx, ny = p->ny;
  dg_basis B;
  make_basis(p->degree, &B);
  const int nq = B.nq, np = nq * nq;
  double lx[3], ly[3];
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i)
      for (int b = 0; b < nq; ++b)
        for (int a = 0; a < nq; ++a) {
          for (int n = 0; n <= B.k; ++n) { lx[n] = B.phi[n][a]; ly[n] = B.phi[n][b]; }
          double dv[4];
          eval_dev(d, i, j, nx, ny, &B, lx, ly, dv);
          for (int v = 0; v < 4; ++v)
            nodal[(size_t)v + 4u * ((size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)(a + nq * b)))] = dv[v];
        }
  (void)np;
}

/* L2 projection of nodal values (layout as orc_dg2d_nodal) onto the modes */
void orc_dg2d_project(const fvwb_dg_params* p, const double* nodal, double* d)
{
  const int nx = p->nx, ny = p->ny;
  dg_basis B;
  make_basis(p->degree, &B);
  const int nq = B.nq;
  memset(d, 0, sizeof(double) * 4 * (size_t)nx * ny * B.nm);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i)
      for (int b = 0; b < nq; ++b)
        for (int a = 0; a < nq; ++a) {
          const double w = 0.25 * B.wq[a] * B.wq[b];
          for (int m = 0; m < B.nm; ++m) {
            const double ph = B.phi[B.mi[m]][a] * B.phi[B.mj[m]][b];
            for (int v = 0; v < 4; ++v)
              DM(d, v, i, j, m, nx, ny) +=
                  w * ph * nodal[(size_t)v + 4u * ((size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)(a + nq * b)))];
          }
        }
}
