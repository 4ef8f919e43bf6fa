// fvwb_dgq.hip — the reference's own 2D modal DG program (2d/benchmark_2d_dg.f90
// with 2d/legendre.f90 and 2d/limiters.f90) on MI355X: the tensor-product (Q)
// Legendre basis of mx = my = O modes per direction, in FP64, equal to the
// reference bit for bit.
//
// What the reference evaluates with libm (the Gauss-Legendre rule by Newton's
// method, the basis values) is computed once on the host with the reference's
// expressions and handed to the kernels as a table; on the device there are
// only +, -, *, /, sqrt, min and max, written in the reference's order with
// contraction off, so they round as the reference's do.
//
// One residual evaluation (compute_update, :1137-1479) is three launches:
//   dgq_trace_kernel  one thread per cell: the traces at xi = -1, +1 and
//                     eta = -1, +1 (:1253-1314) and their fluxes
//                     (compute_flux_int :946-965), written face-major so that
//                     the face pass reads them coalesced;
//   dgq_face_kernel   one thread per face point: compute_num_flux (:991-1006)
//                     between the traces of the two cells, with the
//                     reference's get_boundary_conditions(., 2) indexing;
//   dgq_cell_kernel   one thread per cell: node values, volume integrals,
//                     edge integrals, source, assembly (:1448-1466),
//                     special_boundary_conditions, and, for the resident time
//                     loop, the Runge-Kutta stage combination of evolve in
//                     the reference's expression order.
// dgq_limit_kernel is compute_positivity (2d/limiters.f90:478-654) per cell;
// the dgq_speed_*_kernels are compute_max_speed (:826-870) with its
// bookkeeping and evolve's dt (:671), as exact block reductions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <initializer_list>
#include <mutex>
#include <new>
#include <string>
#include <utility>

#include "../../include/fvwb.h"

#pragma clang fp contract(off)

namespace dgq {

constexpr int MAXO = 3;

// Basis and quadrature of one set-up (order O = mx = my).
struct Tab {
  double xq[MAXO], wq[MAXO];     // gl_quadrature(x_quad, w_quad, O)
  double L[MAXO][MAXO];          // L[a][n] = legendre(x_quad(a), n)
  double P[MAXO][MAXO];          // P[a][n] = legendre_prime(x_quad(a), n)
  double lm[MAXO], lp[MAXO];     // legendre(-1, n), legendre(+1, n)
  double Gl[MAXO][MAXO];         // Gl[c][n] = legendre(x_gll(c), n), compute_set's points
  int gll;
};

struct Geo {
  int nx, ny;
  int flux, source, gpc, bc, ninit, use_limiter;
  double gamma, boxlen_x, boxlen_y, odx, dx, eps, cfl, cmin;   // cmin = min(1/9, gll_w_1/2)
};

// The resident time loop's clock: dt of the current step, whether it runs, totals.
struct Clock {
  double t, dt, dt_last;
  long long steps;
  int active, err;
};

// One Runge-Kutta stage of evolve: out = ((a0 X0 + a1 X1) + ...) + (cdt dt) L,
// the reference's left-to-right order; nterms = 0 writes L itself.  With part
// set, the stage also stores part = (b0 Y0 + b1 Y1) + ... (npart terms: the
// leading terms of a later stage that takes the same L) and lval = L, so that
// the later stage finishes from them instead of evaluating L again.
struct Epi {
  const double* X[4];
  double a[4];
  double cdt;
  int nterms;
  const double* Y[3];
  double b[3];
  int npart;
  double* part;
  double* lval;
};

// ---- host: 2d/legendre.f90 as the reference evaluates it ------------------------
static double h_legendre(double x, int n) {
  x = std::min(std::max(x, -1.0), 1.0);
  const double x2 = x * x;
  double l = 0.0;
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
  return std::sqrt(2.0 * (double)n + 1.0) * l;
}
static double h_legendre_prime(double x, int n) {
  x = std::min(std::max(x, -1.0), 1.0);
  const double x2 = x * x;
  double l = 0.0;
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
  return std::sqrt(2.0 * (double)n + 1.0) * l;
}
// gl_quadrature (:77-108): the initial guess's coefficient in default REAL,
// 500 Newton steps, nodes sign-flipped, weights 2(2n+1)/((1-x^2) L'(x)^2),
// the second half mirrored.
static void h_gl_quadrature(int n, double* xq, double* wq) {
  const double dpi = std::acos(-1.0);
  for (int i = 1; i <= n; ++i) {
    const float nf = (float)n;
    const float c = (1.0f - 0.125f / nf / nf) + 0.125f / nf / nf / nf;
    double xx = (double)c * std::cos(dpi * (4.0 * (double)i - 1.0) / (4.0 * (double)n + 2.0));
    for (int it = 0; it < 500; ++it) xx = xx - h_legendre(xx, n) / h_legendre_prime(xx, n);
    xq[i - 1] = -xx;
    const double lp = h_legendre_prime(xq[i - 1], n);
    wq[i - 1] = 2 * (2.0 * (double)n + 1.0) / (1.0 - xq[i - 1] * xq[i - 1]) / (lp * lp);
  }
  for (int i = n / 2 + 1; i <= n; ++i) {
    xq[i - 1] = -xq[n - i];
    wq[i - 1] = wq[n - i];
  }
}

static Tab make_tab(int o) {
  Tab T{};
  h_gl_quadrature(o, T.xq, T.wq);
  for (int n = 0; n < o; ++n) {
    T.lm[n] = h_legendre(-1.0, n);
    T.lp[n] = h_legendre(1.0, n);
    for (int a = 0; a < o; ++a) {
      T.L[a][n] = h_legendre(T.xq[a], n);
      T.P[a][n] = h_legendre_prime(T.xq[a], n);
    }
  }
  // compute_positivity's point set uses gll_quadrature's points: (-1, 0) for
  // two points and (-1, 0, 1) for three (2d/legendre.f90:111-170 as written)
  T.gll = (2 * (o - 1) + 3) / 2;
  const double xg[3] = {-1.0, 0.0, 1.0};
  for (int c = 0; c < T.gll; ++c)
    for (int n = 0; n < o; ++n) T.Gl[c][n] = h_legendre(xg[c], n);
  return T;
}

// ---- device: the reference's point functions ------------------------------------
__device__ __forceinline__ double fmax_r(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double fmin_r(double a, double b) { return a < b ? a : b; }

// compute_primitive (:891-902): density floored at the default-real 10e-10
__device__ __forceinline__ void prim(double gamma, const double u[4], double w[4]) {
  const double fl = (double)10e-10f;
  w[0] = u[0] > fl ? u[0] : fl;
  w[1] = u[1] / w[0];
  w[2] = u[2] / w[0];
  w[3] = (gamma - 1.0) * (u[3] - 0.5 * w[0] * (w[1] * w[1] + w[2] * w[2]));
}

// compute_flux (:919-944): f1 along x, f2 along y
__device__ __forceinline__ void flux_full(double gamma, const double u[4], double f1[4], double f2[4]) {
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

// compute_flux_int (:946-965) in one direction: dir 0 along x, 1 along y
__device__ __forceinline__ void flux_int(double gamma, const double u[4], int dir, double f[4]) {
  double w[4];
  prim(gamma, u, w);
  if (dir == 0) {
    f[0] = w[1] * u[0];
    f[1] = w[1] * u[1] + w[3];
    f[2] = w[0] * w[1] * w[2];
    f[3] = w[1] * u[3] + w[1] * w[3];
  } else {
    f[0] = w[2] * u[0];
    f[1] = w[0] * w[1] * w[2];
    f[2] = w[2] * u[2] + w[3];
    f[3] = w[2] * u[3] + w[2] * w[3];
  }
}

// compute_speed (:873-889)
__device__ __forceinline__ void speed(double gamma, const double u[4], double& cs, double& vx, double& vy,
                                      double& sp) {
  double w[4];
  prim(gamma, u, w);
  cs = sqrt(gamma * fmax(w[3], 1e-10) / fmax(w[0], 1e-10));
  vx = w[1];
  vy = w[2];
  sp = sqrt(w[1] * w[1] + w[2] * w[2]) + cs;
}

// compute_num_flux (:991-1006); flag 1 along x, 2 along y.  A flux_type of no
// branch (the shipped 'llf') and an HLLC state of no branch leave out at zero,
// the value the reference's F and G start from (:1317-1318).
__device__ void num_flux(const Geo& g, const double ul[4], const double ur[4], const double fl[4],
                         const double fr[4], int flag, double out[4]) {
  out[0] = out[1] = out[2] = out[3] = 0.0;
  const double gamma = g.gamma;
  if (g.flux == FVWB_DGQ_FLUX_LLF1) {            // compute_llflux (:968-988)
    double csl, vxl, vyl, spl, csr, vxr, vyr, spr;
    speed(gamma, ul, csl, vxl, vyl, spl);
    speed(gamma, ur, csr, vxr, vyr, spr);
    const double cmax = flag == 1 ? fmax_r(fabs(vxr + csr), fabs(vxl + csl)) : fmax_r(fabs(vyr + csr), fabs(vyl + csl));
#pragma unroll
    for (int v = 0; v < 4; ++v) out[v] = 0.5 * (fr[v] + fl[v]) + 0.5 * cmax * (ul[v] - ur[v]);
  } else if (g.flux == FVWB_DGQ_FLUX_HLL2) {     // compute_hllflux (:1008-1026)
    double csl, vxl, vyl, spl, csr, vxr, vyr, spr;
    speed(gamma, ul, csl, vxl, vyl, spl);
    speed(gamma, ur, csr, vxr, vyr, spr);
    const double ap = fmax_r(0.0, fmax_r(csl + sqrt(vxl * vxl + vyl * vyl), csr + sqrt(vxr * vxr + vyr * vyr)));
    const double am = fmax_r(0.0, fmax_r(-(csl - sqrt(vxl * vxl + vyl * vyl)), -(csr - sqrt(vxr * vxr + vyr * vyr))));
#pragma unroll
    for (int v = 0; v < 4; ++v) out[v] = (ap * fl[v] + am * fr[v] - ap * am * (ur[v] - ul[v])) / (ap + am);
  } else if (g.flux == FVWB_DGQ_FLUX_HLLC) {     // compute_hllcflux (:1030-1134) as written
    double wl[4], wr[4], csl, vxl, vyl, spl, csr, vxr, vyr, spr;
    prim(gamma, ul, wl);
    prim(gamma, ur, wr);
    speed(gamma, ul, csl, vxl, vyl, spl);
    speed(gamma, ur, csr, vxr, vyr, spr);
    const double vl = flag == 1 ? vxl : vyl, vr = flag == 1 ? vxr : vyr;
    const double SL = fmin_r(vl, vr) - fmax_r(csl, csr);
    const double SR = fmax_r(vl, vr) + fmax_r(csl, csr);
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
      usr[1] = usr[0] * wl[1];                   // the left tangential velocity, as written
      usr[2] = usr[0] * SM;
      usr[3] = usr[0] * (ur[3] / ur[0] + (SM - wr[2] * (SM + wr[3] / (wr[0] * (SR - wr[2])))));
    }
    double f1[4], f2[4];
    int branch = 0;
    if (SL > 0.0) branch = 1;
    else if (SL <= 0.0 && SM > 0.0) branch = 2;
    else if (SR >= 0.0 && SM <= 0.0) branch = 3;
    else if (SR < 0.0) branch = 4;
    if (branch) {
      flux_full(gamma, branch <= 2 ? ul : ur, f1, f2);
      const double* f = flag == 1 ? f1 : f2;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        if (branch == 1 || branch == 4) out[v] = f[v];
        else if (branch == 2) out[v] = f[v] + SL * (usl[v] - ul[v]);
        else out[v] = f[v] + SR * (usr[v] - ur[v]);
      }
    }
  }
}

// get_boundary_conditions(index, 2) (:777-824) with one-based indices; the
// reference passes dim = 2 for the x faces as well (:1338-1339), which is why
// the residual needs nx == ny
__device__ __forceinline__ int bc_index(int index, int n, int bc) {
  if (bc == 1) {
    if (index == 0) return n;
    if (index == n + 1) return 1;
  } else if (bc == 2 || bc == 3) {
    if (index == 0) return 1;
    if (index == n + 1) return n;
  }
  return index;
}

// grad_phi (:1599-1644) at one point
__device__ __forceinline__ void grad_phi(int gcase, double x, double y, double& g0, double& g1) {
  if (gcase == 1) {
    g0 = x;
    g1 = y;
  } else if (gcase == 2) {
    const double epsilon = 0.25, delta_r = (double)0.1f, xd = x - 3., yd = y - 3.;
    const double r = sqrt(xd * xd + yd * yd);
    if (r > 0.5 - 0.5 * delta_r) {
      const double r3 = r * r * r;
      g0 = -(xd) / r3;
      g1 = -(yd) / r3;
    } else {
      g0 = -(xd) / (r * (r * r + epsilon * epsilon));
      g1 = -(yd) / (r * (r * r + epsilon * epsilon));
    }
  } else {
    g0 = g1 = 0.0;
  }
}

// ---- kernels ------------------------------------------------------------------
// Element (v, cell c, a, b) of a modal or nodal array at v + 4*(c + nc*(a + O*b)).
template <int O>
__device__ __forceinline__ long long at(int v, long long c, long long nc, int a, int b) {
  return (long long)v + 4 * (c + nc * ((long long)a + O * b));
}

// Trace buffer: states then fluxes, element ((side*O + q)*4 + v)*nc + c with
// side 0 = left (xi = -1), 1 = right, 2 = bottom (eta = -1), 3 = top; the
// flux of a left/right trace is the x flux, of a bottom/top trace the y flux.
template <int O>
__global__ __launch_bounds__(256) void dgq_trace_kernel(Geo g, Tab T, const double* __restrict__ du,
                                                        double* __restrict__ tr, const Clock* __restrict__ ck) {
  if (ck && !ck->active) return;
  const long long nc = (long long)g.nx * g.ny;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  double m[4][O][O];
#pragma unroll
  for (int i = 0; i < O; ++i)
#pragma unroll
    for (int j = 0; j < O; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) m[v][i][j] = du[at<O>(v, c, nc, i, j)];
  double* fl = tr + 16LL * O * nc;
#pragma unroll
  for (int side = 0; side < 4; ++side) {
#pragma unroll
    for (int q = 0; q < O; ++q) {
      double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int i = 0; i < O; ++i)
#pragma unroll
        for (int j = 0; j < O; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            // :1259-1266 (x sides) and :1288-1297 (y sides)
            const double b1 = side == 0 ? T.lm[i] : side == 1 ? T.lp[i] : side == 2 ? T.lm[j] : T.lp[j];
            const double b2 = side < 2 ? T.L[q][j] : T.L[q][i];
            s[v] = s[v] + m[v][i][j] * b1 * b2;
          }
      double f[4];
      flux_int(g.gamma, s, side < 2 ? 0 : 1, f);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        tr[((long long)(side * O + q) * 4 + v) * nc + c] = s[v];
        fl[((long long)(side * O + q) * 4 + v) * nc + c] = f[v];
      }
    }
  }
}

// Face fluxes: x faces F(v, q, f, j) at ((f + (nx+1) j) O + q) 4 + v, f = 0..nx;
// y faces G(v, q, i, f) at ((i + nx f) O + q) 4 + v, f = 0..ny.
template <int O>
__global__ __launch_bounds__(256) void dgq_face_kernel(Geo g, const double* __restrict__ tr,
                                                       double* __restrict__ F, double* __restrict__ G,
                                                       const Clock* __restrict__ ck) {
  if (ck && !ck->active) return;
  const long long nc = (long long)g.nx * g.ny;
  const long long nfx = (long long)(g.nx + 1) * g.ny * O, nfy = (long long)g.nx * (g.ny + 1) * O;
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nfx + nfy) return;
  const double* fl = tr + 16LL * O * nc;
  const bool xdir = k < nfx;
  if (!xdir) k -= nfx;
  const int q = (int)(k % O);
  const long long r = k / O;
  long long cl, cr;
  int sl, sr;
  if (xdir) {
    const int f = (int)(r % (g.nx + 1)), j = (int)(r / (g.nx + 1));
    const int il = bc_index(f, g.ny, g.bc) - 1, ir = bc_index(f + 1, g.ny, g.bc) - 1;
    cl = il + (long long)g.nx * j;       // right trace of the left cell
    cr = ir + (long long)g.nx * j;       // left trace of the right cell
    sl = 1;
    sr = 0;
  } else {
    const int i = (int)(r % g.nx), f = (int)(r / g.nx);
    const int jl = bc_index(f, g.ny, g.bc) - 1, jr = bc_index(f + 1, g.ny, g.bc) - 1;
    cl = i + (long long)g.nx * jl;       // top trace of the cell below
    cr = i + (long long)g.nx * jr;       // bottom trace of the cell above
    sl = 3;
    sr = 2;
  }
  double ul[4], ur[4], ful[4], fur[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    ul[v] = tr[((long long)(sl * O + q) * 4 + v) * nc + cl];
    ur[v] = tr[((long long)(sr * O + q) * 4 + v) * nc + cr];
    ful[v] = fl[((long long)(sl * O + q) * 4 + v) * nc + cl];
    fur[v] = fl[((long long)(sr * O + q) * 4 + v) * nc + cr];
  }
  double out[4];
  num_flux(g, ul, ur, ful, fur, xdir ? 1 : 2, out);
  double* dst = (xdir ? F : G) + 4 * k;
#pragma unroll
  for (int v = 0; v < 4; ++v) dst[v] = out[v];
}

// The residual of one cell, one variable at a time (the node values and
// fluxes are recomputed per variable so the accumulators of one variable fit
// in registers), then the stage combination.  Accumulation orders are the
// reference's: over the nodes a outer, b inner for the volume and source
// integrals, over the face points for the edges.  The residual of the
// block's cells waits in LDS until all four variables are done: for a node
// (i, j) the block's cells and variables are one contiguous run of the
// output (element v + 4 (c + nc (i + O j))), which the block then combines
// and stores in full cache lines.
constexpr int CELL_THREADS = 128;

template <int O>
__global__ __launch_bounds__(CELL_THREADS) __attribute__((amdgpu_waves_per_eu(2))) void dgq_cell_kernel(
    Geo g, Tab T, const double* __restrict__ in, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ F, const double* __restrict__ G, Epi e, double* out, const Clock* __restrict__ ck) {
  if (ck && !ck->active) return;
  __shared__ double sres[O * O * CELL_THREADS * 4];
  const long long nc = (long long)g.nx * g.ny;
  const long long c0 = (long long)blockIdx.x * CELL_THREADS;
  const int tid = threadIdx.x;
  const long long c = c0 + tid;
  const double dt = ck ? ck->dt : 0.0;
  if (c < nc) {
    const int ic = (int)(c % g.nx), jc = (int)(c / g.nx);
    double m[4][O][O];
#pragma unroll
    for (int i = 0; i < O; ++i)
#pragma unroll
      for (int j = 0; j < O; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) m[v][i][j] = in[at<O>(v, c, nc, i, j)];
    const double odx = g.odx;
    // two variables per pass: the nodal state and its fluxes, which every
    // variable needs, are formed once for the pair; each accumulator still
    // takes its terms in the reference's order
#pragma unroll 1
    for (int vp = 0; vp < 4; vp += 2) {
      double fv1[2][O][O], fv2[2][O][O], sv[2][O][O];
#pragma unroll
      for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int i = 0; i < O; ++i)
#pragma unroll
          for (int j = 0; j < O; ++j) fv1[r][i][j] = fv2[r][i][j] = sv[r][i][j] = 0.0;
      // the node loops stay rolled: unrolled, the kernel held every table entry
      // and node value at once (496 VGPRs, one wave per SIMD); the order of the
      // sums is the same either way
#pragma unroll 1
      for (int a = 0; a < O; ++a)
#pragma unroll 1
        for (int b = 0; b < O; ++b) {
          // get_nodes_from_modes (:544-592): modes a' outer, b' inner
          double u[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int ma = 0; ma < O; ++ma)
#pragma unroll
            for (int mb = 0; mb < O; ++mb)
#pragma unroll
              for (int w = 0; w < 4; ++w) u[w] = u[w] + m[w][ma][mb] * T.L[a][ma] * T.L[b][mb];
          double f1[4], f2[4];
          flux_full(g.gamma, u, f1, f2);
          double s[2] = {0.0, 0.0};
          bool has_s = false;
          if (g.source == 2) {                     // get_source (:1558-1576)
            double w[4], g0, g1;
            prim(g.gamma, u, w);
            grad_phi(g.gpc, x[c + nc * (a + O * b)], y[c + nc * (a + O * b)], g0, g1);
#pragma unroll
            for (int r = 0; r < 2; ++r) {
              const int v = vp + r;
              s[r] = v == 0 ? 0. : v == 1 ? w[0] * g0 : v == 2 ? w[0] * g1 : w[0] * (w[1] * g0 + w[2] * g1);
            }
            has_s = true;
          } else if (g.source == 3) {              // get_adv_source (:1579-1596)
#pragma unroll
            for (int r = 0; r < 2; ++r) s[r] = vp + r == 0 ? -1.0 * u[0] : 0.0;
            has_s = true;
          }
#pragma unroll
          for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int i = 0; i < O; ++i)
#pragma unroll
              for (int j = 0; j < O; ++j) {
                fv1[r][i][j] = fv1[r][i][j] + f1[vp + r] * T.P[a][i] * T.wq[a] * T.L[b][j] * T.wq[b];
                fv2[r][i][j] = fv2[r][i][j] + f2[vp + r] * T.P[b][j] * T.wq[b] * T.L[a][i] * T.wq[a];
                if (has_s) sv[r][i][j] = sv[r][i][j] + s[r] * T.L[a][i] * T.wq[a] * T.L[b][j] * T.wq[b];
              }
        }
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const int v = vp + r;
        // face values of this variable: F at the cell's left (ic) and right
        // (ic+1) x faces, G at its bottom (jc) and top (jc+1) y faces
        double fr[O], fle[O], gt[O], gb[O];
#pragma unroll
        for (int q = 0; q < O; ++q) {
          fle[q] = F[4 * (((long long)ic + (long long)(g.nx + 1) * jc) * O + q) + v];
          fr[q] = F[4 * (((long long)ic + 1 + (long long)(g.nx + 1) * jc) * O + q) + v];
          gb[q] = G[4 * (((long long)ic + (long long)g.nx * jc) * O + q) + v];
          gt[q] = G[4 * (((long long)ic + (long long)g.nx * (jc + 1)) * O + q) + v];
        }
#pragma unroll
        for (int i = 0; i < O; ++i)
#pragma unroll
          for (int j = 0; j < O; ++j) {
            double e1 = 0.0, e2 = 0.0, e3 = 0.0, e4 = 0.0;
#pragma unroll
            for (int q = 0; q < O; ++q) {          // :1377-1386
              e1 = e1 + fr[q] * T.lp[i] * T.L[q][j] * T.wq[q];
              e2 = e2 + fle[q] * T.lm[i] * T.L[q][j] * T.wq[q];
            }
#pragma unroll
            for (int q = 0; q < O; ++q) {          // :1391-1408
              e3 = e3 + gt[q] * T.lp[j] * T.L[q][i] * T.wq[q];
              e4 = e4 + gb[q] * T.lm[j] * T.L[q][i] * T.wq[q];
            }
            // :1455-1462 (oneoverdx for the y terms too)
            double d = (odx * fv1[r][i][j] + odx * fv2[r][i][j] - odx * (e1 - e2) - odx * (e3 - e4)) / 2. +
                       sv[r][i][j] / 4.;
            if (g.ninit == 12) {                   // special_boundary_conditions (:1481-1514)
              const double xd = x[c + nc * (i + O * j)] - g.boxlen_x / 2.;
              const double yd = y[c + nc * (i + O * j)] - g.boxlen_y / 2.;
              if (sqrt(xd * xd + yd * yd) > 2.0) d = 0.0;
            }
            sres[((i + O * j) * CELL_THREADS + tid) * 4 + v] = d;
          }
      }
    }
  }
  __syncthreads();
  // node (i, j): elements [4 (c0 + nc (i + O j)), + 4 ncells) of every array
  const int ncell = (int)(nc - c0 < CELL_THREADS ? nc - c0 : CELL_THREADS);
#pragma unroll 1
  for (int ij = 0; ij < O * O; ++ij) {
    const long long base = 4 * (c0 + nc * ij);
    for (int q = tid; q < 4 * ncell; q += CELL_THREADS) {
      const long long idx = base + q;
      const double d = sres[ij * CELL_THREADS * 4 + q];
      if (e.nterms == 0) {
        out[idx] = d;
      } else {
        double sacc = e.a[0] * e.X[0][idx];
        for (int k = 1; k < e.nterms; ++k) sacc = sacc + e.a[k] * e.X[k][idx];
        out[idx] = sacc + e.cdt * dt * d;
        if (e.part) {
          double pacc = e.b[0] * e.Y[0][idx];
          for (int k = 1; k < e.npart; ++k) pacc = pacc + e.b[k] * e.Y[k][idx];
          e.part[idx] = pacc;
          e.lval[idx] = d;
        }
      }
    }
  }
}

// The rest of a stage whose leading terms and L a previous stage stored (Epi
// part / lval): out = (part + a X) + (cdt dt) L, the same operations in the
// same order as the stage evaluated in one piece.  out may alias part.
__global__ __launch_bounds__(256) void dgq_finish_kernel(long long n, const double* part, double a,
                                                         const double* __restrict__ X, double cdt,
                                                         const double* __restrict__ lval, double* out,
                                                         const Clock* __restrict__ ck) {
  if (!ck->active) return;
  const double dt = ck->dt;
  for (long long idx = (long long)blockIdx.x * 256 + threadIdx.x; idx < n; idx += (long long)gridDim.x * 256) {
    double sacc = part[idx];
    sacc = sacc + a * X[idx];
    out[idx] = sacc + cdt * dt * lval[idx];
  }
}

// solve_for_t (2d/limiters.f90:312-362): the root in [0,1] of the quadratic
// whose zero puts the pressure of (1-t) u_avg + t u at eps
__device__ double solve_for_t(double gamma, double eps, const double u[4], const double ua[4]) {
  const double pa = ua[0], mxa = ua[1], mya = ua[2], ea = ua[3];
  const double pj = u[0], mxj = u[1], myj = u[2], ej = u[3];
  const double a = 2.0 * (pj - pa) * (ej - ea) - (mxj - mxa) * (mxj - mxa) - (myj - mya) * (myj - mya);
  double b = 2.0 * (pj - pa) * (ea - eps / (gamma - 1)) + 2.0 * pa * (ej - ea) -
             2.0 * (mxa * (mxj - mxa) + mya * (myj - mya));
  double cc = 2.0 * pa * ea - (mxa * mxa + mya * mya) - 2.0 * eps * pa / (gamma - 1.0);
  b = b / a;
  cc = cc / a;
  const double D = sqrt(fabs(b * b - 4 * cc));
  const double t1 = 0.5 * (-b - D), t2 = 0.5 * (-b + D);
  double t;
  if (t1 > -eps && t1 < 1.0 + eps) t = t1;
  else if (t2 > -eps && t2 < 1.0 + eps) t = t2;
  else t = 0.0;
  t = fmin(1.0, t);
  t = fmax(0.0, t);
  return t;
}

// compute_set (2d/limiters.f90:438-475), point k of 2 O gll: the first O gll
// points at (x_gll(c), y_gl(q)), k = q gll + c, the others at (x_gl(q), y_gll(c))
template <int O>
__device__ __forceinline__ void set_point(const Tab& T, const double m[4][O][O], int k, double p[4]) {
  const int half = O * T.gll;
  const bool second = k >= half;
  const int kk = second ? k - half : k;
  const int q = kk / T.gll, cg = kk % T.gll;
  p[0] = p[1] = p[2] = p[3] = 0.0;
#pragma unroll
  for (int i = 0; i < O; ++i)
#pragma unroll
    for (int j = 0; j < O; ++j) {
      const double bi = second ? T.L[q][i] : T.Gl[cg][i];
      const double bj = second ? T.Gl[cg][j] : T.L[q][j];
#pragma unroll
      for (int v = 0; v < 4; ++v) p[v] = p[v] + m[v][i][j] * bi * bj;
    }
}

// compute_positivity (2d/limiters.f90:478-654) on one cell, in place: theta
// over the set's densities scales the non-mean density modes, then the
// smallest solve_for_t over the set points whose pressure (density floored at
// the default-real 10e-10) is not above eps scales every non-mean mode.
template <int O>
__global__ __launch_bounds__(256) void dgq_limit_kernel(Geo g, Tab T, double* __restrict__ u,
                                                        const Clock* __restrict__ ck) {
  if (ck && !ck->active) return;
  const long long nc = (long long)g.nx * g.ny;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  double m[4][O][O];
#pragma unroll
  for (int i = 0; i < O; ++i)
#pragma unroll
    for (int j = 0; j < O; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) m[v][i][j] = u[at<O>(v, c, nc, i, j)];
  const int np = 2 * O * T.gll;
  const double eps = g.eps, gamma = g.gamma;
  {
    double pmin = 0.0;
    for (int k = 0; k < np; ++k) {
      double p[4];
      set_point<O>(T, m, k, p);
      pmin = k == 0 ? p[0] : fmin(pmin, p[0]);
    }
    const double theta = fmin(fabs((m[0][0][0] - eps) / (m[0][0][0] - pmin)), 1.0);
#pragma unroll
    for (int i = 0; i < O; ++i)
#pragma unroll
      for (int j = 0; j < O; ++j)
        if (i != 0 || j != 0) m[0][i][j] = theta * m[0][i][j];
  }
  const double avg[4] = {m[0][0][0], m[1][0][0], m[2][0][0], m[3][0][0]};
  double tmin = 1.;
  for (int k = 0; k < np; ++k) {
    double p[4];
    set_point<O>(T, m, k, p);
    const double r = fmax(p[0], (double)10e-10f);
    const double vx = p[1] / r, vy = p[2] / r;
    const double pr = (gamma - 1.0) * (p[3] - 0.5 * r * (vx * vx + vy * vy));
    const double t = pr > eps ? 1. : solve_for_t(gamma, eps, p, avg);
    if (tmin >= t) tmin = t;
  }
#pragma unroll
  for (int i = 0; i < O; ++i)
#pragma unroll
    for (int j = 0; j < O; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double val = (i != 0 || j != 0) ? tmin * m[v][i][j] : m[v][i][j];
        u[at<O>(v, c, nc, i, j)] = val;
      }
}

// get_nodes_from_modes / get_modes_from_nodes (:497-592) per cell, in the
// reference's loop order (modes a' outer, b' inner; for the modes the nodes
// a outer, b inner with 0.25 first)
template <int O, bool TO_NODES>
__global__ __launch_bounds__(256) void dgq_project_kernel(Geo g, Tab T, const double* __restrict__ src,
                                                          double* __restrict__ dst) {
  const long long nc = (long long)g.nx * g.ny;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  double s[4][O][O];
#pragma unroll
  for (int i = 0; i < O; ++i)
#pragma unroll
    for (int j = 0; j < O; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) s[v][i][j] = src[at<O>(v, c, nc, i, j)];
#pragma unroll
  for (int i = 0; i < O; ++i)
#pragma unroll
    for (int j = 0; j < O; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        double acc = 0.0;
#pragma unroll
        for (int a = 0; a < O; ++a)
#pragma unroll
          for (int b = 0; b < O; ++b) {
            if (TO_NODES)
              acc = acc + s[v][a][b] * T.L[i][a] * T.L[j][b];
            else
              acc = acc + 0.25 * s[v][a][b] * T.L[a][i] * T.L[b][j] * T.wq[a] * T.wq[b];
          }
        dst[at<O>(v, c, nc, i, j)] = acc;
      }
}

// compute_max_speed (:826-870) over the cells in the reference's order
// k = i ny + j (x outer) and, with a clock, evolve's dt (:668-671).  The
// maximum speed S, the last k whose speed equals S (the last update of the
// reference's running ">=" test) and the smallest sound speed from that cell
// on (cs_max restarts there and only decreases) are each exact reductions, so
// their order does not matter.  Three launches: per block the pair (S_b, last
// k with speed S_b); per block the smallest sound speed at k >= k* (every
// block first combines the pairs itself); one block that combines everything,
// re-evaluates the cell k* and sets dt.  res = (cs_max, v_xmax, v_ymax,
// speed_max); part holds three doubles per block.
constexpr int SPEED_THREADS = 256;
constexpr int SPEED_MAX_BLOCKS = 1024;

__device__ __forceinline__ void cell_speed(const Geo& g, const double* __restrict__ u, long long m, double& cs,
                                           double& vx, double& vy, double& sp) {
  double q[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) q[v] = u[4 * m + v];
  speed(g.gamma, q, cs, vx, vy, sp);
}

// (s, k) pairs: the larger speed wins, equal speeds keep the larger k
__device__ __forceinline__ void pair_max(double& s, long long& k, double so, long long ko) {
  if (so > s) { s = so; k = ko; }
  else if (so == s && ko > k) k = ko;
}

__device__ __forceinline__ void wave_pair_max(double& s, long long& k) {
  for (int off = 32; off > 0; off >>= 1) {
    const double so = __shfl_xor(s, off);
    const long long ko = __shfl_xor(k, off);
    pair_max(s, k, so, ko);
  }
}

// a step that will not run (clock stopped or past t_end) skips the passes
__device__ __forceinline__ bool speed_runs(const Clock* ck, double t_end) {
  return !ck || (!ck->err && ck->t < t_end);
}

// the pairs of every block combined (each thread of the block strides over
// them, then the block reduces): S and k* (-1 when no speed reached S >= 0)
__device__ void combine_pairs(const double* __restrict__ part, int nblk, double& S, long long& ks) {
  __shared__ double sd[SPEED_THREADS / 64];
  __shared__ long long sk[SPEED_THREADS / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double s = 0.0;
  long long k = -1;
  for (int b = tid; b < nblk; b += SPEED_THREADS) pair_max(s, k, part[3 * b], (long long)part[3 * b + 1]);
  wave_pair_max(s, k);
  if (lane == 0) { sd[wv] = s; sk[wv] = k; }
  __syncthreads();
  s = sd[0];
  k = sk[0];
  for (int w = 1; w < SPEED_THREADS / 64; ++w) pair_max(s, k, sd[w], sk[w]);
  S = s;
  ks = k;
  __syncthreads();
}

__global__ __launch_bounds__(SPEED_THREADS) void dgq_speed_pairs_kernel(Geo g, const double* __restrict__ u,
                                                                       double* __restrict__ part,
                                                                       const Clock* __restrict__ ck, double t_end) {
  if (!speed_runs(ck, t_end)) return;
  __shared__ double sd[SPEED_THREADS / 64];
  __shared__ long long sk[SPEED_THREADS / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const long long nc = (long long)g.nx * g.ny;
  // speeds start at 0 with k = -1: a speed that is 0 joins the pair at 0, a
  // NaN never compares, as in the reference's running test from 0
  double s = 0.0;
  long long k = -1;
  for (long long m = (long long)blockIdx.x * SPEED_THREADS + tid; m < nc; m += (long long)gridDim.x * SPEED_THREADS) {
    double cs, vx, vy, sp;
    cell_speed(g, u, m, cs, vx, vy, sp);
    const long long km = (m % g.nx) * g.ny + m / g.nx;
    if (sp > s) { s = sp; k = km; }
    else if (sp == s && km > k) k = km;
  }
  wave_pair_max(s, k);
  if (lane == 0) { sd[wv] = s; sk[wv] = k; }
  __syncthreads();
  if (tid == 0) {
    for (int w = 1; w < SPEED_THREADS / 64; ++w) pair_max(s, k, sd[w], sk[w]);
    part[3 * blockIdx.x] = s;
    part[3 * blockIdx.x + 1] = (double)k;   // exact: k < 2^53
  }
}

__global__ __launch_bounds__(SPEED_THREADS) void dgq_speed_csmin_kernel(Geo g, const double* __restrict__ u,
                                                                       double* __restrict__ part,
                                                                       const Clock* __restrict__ ck, double t_end) {
  if (!speed_runs(ck, t_end)) return;
  __shared__ double sd[SPEED_THREADS / 64];
  double S;
  long long ks;
  combine_pairs(part, gridDim.x, S, ks);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const long long nc = (long long)g.nx * g.ny;
  double csm = INFINITY;
  if (ks >= 0)
    for (long long m = (long long)blockIdx.x * SPEED_THREADS + tid; m < nc; m += (long long)gridDim.x * SPEED_THREADS) {
      const long long km = (m % g.nx) * g.ny + m / g.nx;
      if (km < ks) continue;
      double cs, vx, vy, sp;
      cell_speed(g, u, m, cs, vx, vy, sp);
      if (csm > cs) csm = cs;
    }
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(csm, off);
    if (csm > o) csm = o;
  }
  if (lane == 0) sd[wv] = csm;
  __syncthreads();
  if (tid == 0) {
    for (int w = 1; w < SPEED_THREADS / 64; ++w) if (csm > sd[w]) csm = sd[w];
    part[3 * blockIdx.x + 2] = csm;
  }
}

__global__ __launch_bounds__(SPEED_THREADS) void dgq_speed_final_kernel(Geo g, const double* __restrict__ u,
                                                                       const double* __restrict__ part, int nblk,
                                                                       double* __restrict__ res,
                                                                       Clock* __restrict__ ck, double t_end) {
  if (!speed_runs(ck, t_end)) {
    if (threadIdx.x == 0) ck->active = 0;
    return;
  }
  double S;
  long long ks;
  combine_pairs(part, nblk, S, ks);
  // the smallest of the blocks' sound-speed minima, in parallel
  __shared__ double sc[SPEED_THREADS / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double cmin = INFINITY;
  for (int b = tid; b < nblk; b += SPEED_THREADS) if (cmin > part[3 * b + 2]) cmin = part[3 * b + 2];
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(cmin, off);
    if (cmin > o) cmin = o;
  }
  if (lane == 0) sc[wv] = cmin;
  __syncthreads();
  if (tid != 0) return;
  for (int w = 1; w < SPEED_THREADS / 64; ++w) if (cmin > sc[w]) cmin = sc[w];
  double cs_max = 0.0, vxm = 0.0, vym = 0.0;
  if (ks >= 0) {
    const long long mb = (ks / g.ny) + (long long)g.nx * (ks % g.ny);
    double sp;
    cell_speed(g, u, mb, cs_max, vxm, vym, sp);
    if (cs_max > cmin) cs_max = cmin;
  }
  if (res) {
    res[0] = cs_max;
    res[1] = vxm;
    res[2] = vym;
    res[3] = S;
  }
  if (ck) {
    const double dtc = g.cfl * g.cmin / ((fabs(vxm) + (cs_max)) / g.dx + (fabs(vym) + (cs_max)) / g.dx);
    const double rest = t_end - ck->t;
    const double dt = rest < dtc ? rest : dtc;
    if (!(dt > 0.0) || !isfinite(dt)) {
      ck->err = 1;
      ck->active = 0;
      return;
    }
    ck->dt = dt;
    ck->active = 1;
  }
}

// t = t + dt (:749), after the step
__global__ void dgq_tick_kernel(Clock* __restrict__ ck) {
  if (threadIdx.x != 0 || !ck->active) return;
  ck->t = ck->t + ck->dt;
  ck->dt_last = ck->dt;
  ck->steps += 1;
}

// compute_error (:23-89) per cell, then a fixed-order sum of the block
// partials: part[blk*12 + (0..3 max, 4..7 l1, 8..11 l2 sums)]
constexpr int ERR_THREADS = 256;

template <int O>
__global__ __launch_bounds__(ERR_THREADS) void dgq_error_kernel(Geo g, Tab T, const double* __restrict__ u,
                                                                const double* __restrict__ u0,
                                                                double* __restrict__ part) {
  __shared__ double sh[ERR_THREADS / 64][12];
  const long long nc = (long long)g.nx * g.ny;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double r[12];
  for (int k = 0; k < 12; ++k) r[k] = 0.0;
  if (c < nc) {
    const double cdx = 1. / (double)g.nx, cdy = 1. / (double)g.ny;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      double a1 = 0.0, a2 = 0.0, mx = 0.0;
#pragma unroll
      for (int a = 0; a < O; ++a)
#pragma unroll
        for (int b = 0; b < O; ++b) {
          const long long idx = at<O>(v, c, nc, a, b);
          const double d = u[idx] - u0[idx];
          mx = fmax(mx, fabs(d));
          a1 = a1 + fabs(d) * T.wq[a] * T.wq[b];
          a2 = a2 + d * d * T.wq[a] * T.wq[b];
        }
      r[v] = mx;
      r[4 + v] = a1 * (cdx * cdy) * 0.25;
      r[8 + v] = a2 * (cdx * cdy) * 0.25;
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int k = 0; k < 12; ++k) {
    double x = r[k];
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_xor(x, off);
      x = k < 4 ? fmax(x, o) : x + o;
    }
    if (lane == 0) sh[wv][k] = x;
  }
  __syncthreads();
  if (threadIdx.x < 12) {
    const int k = threadIdx.x;
    double x = sh[0][k];
    for (int w = 1; w < ERR_THREADS / 64; ++w) x = k < 4 ? fmax(x, sh[w][k]) : x + sh[w][k];
    part[(long long)blockIdx.x * 12 + k] = x;
  }
}

__global__ void dgq_error_final_kernel(const double* __restrict__ part, int nblk, double* __restrict__ res) {
  const int k = threadIdx.x;
  if (k >= 12) return;
  double x = 0.0;
  for (int b = 0; b < nblk; ++b) x = k < 4 ? fmax(x, part[(long long)b * 12 + k]) : x + part[(long long)b * 12 + k];
  res[k] = k >= 8 ? sqrt(x) : x;
}

}  // namespace dgq

using namespace dgq;

struct fvwb_dgq {
  fvwb_dgq_params p{};
  Geo g{};
  Tab T{};
  hipStream_t stream = nullptr;
  long long nc = 0, ns = 0;                  // cells, doubles per state
  double* du = nullptr;                      // resident state
  double* w[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};   // stage states; w[5] holds L(w3)
  double* s_in = nullptr;                    // drop-in staging
  double* s_out = nullptr;
  double *x = nullptr, *y = nullptr;         // resident coordinates
  double *xi = nullptr, *yi = nullptr;       // drop-in coordinates
  double* tr = nullptr;                      // traces and their fluxes
  double *F = nullptr, *G = nullptr;
  double* res = nullptr;                     // small results
  double* part = nullptr;                    // error partials
  Clock* clk = nullptr;
  bool loaded = false;
  std::string err;
};

static thread_local std::string g_errq;
// the message of the allocation-free entry point's last failure (a literal,
// so reporting it allocates nothing); cleared by every other failure
static thread_local const char* g_errq_static = nullptr;

static int failq(fvwb_dgq* h, int code, const std::string& msg) {
  (h ? h->err : g_errq) = msg;
  if (!h) g_errq_static = nullptr;
  return code;
}

#define HIPQ(h, call)                                                                                 \
  do {                                                                                                \
    hipError_t e_ = (call);                                                                           \
    if (e_ != hipSuccess) return failq(h, FVWB_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

extern "C" const char* fvwb_dgq_last_error(fvwb_dgq_handle h) {
  if (h) return h->err.c_str();
  return g_errq_static ? g_errq_static : g_errq.c_str();
}

extern "C" int fvwb_dgq_params_reference(fvwb_dgq_params* p) {
  if (!p) return failq(nullptr, FVWB_ERR_ARG, "NULL argument");
  *p = fvwb_dgq_params{};
  p->nx = p->ny = 8;
  p->order = 2;
  p->flux_type = FVWB_DGQ_FLUX_NONE;
  p->source = 1;
  p->grad_phi_case = 2;
  p->bc = 1;
  p->ninit = 1;
  p->solver = FVWB_DGQ_SOLVER_RK4;
  p->use_limiter = 1;
  p->gamma = (double)1.4f;
  p->boxlen_x = p->boxlen_y = 1.0;
  p->cfl = (double)0.2f;
  p->eps = (double)1e-10f;
  return FVWB_OK;
}

extern "C" int fvwb_dgq_destroy(fvwb_dgq_handle h) {
  if (!h) return FVWB_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  void* ps[] = {h->du, h->w[0], h->w[1], h->w[2], h->w[3], h->w[4], h->w[5], h->s_in, h->s_out, h->x, h->y, h->xi,
                h->yi, h->tr, h->F, h->G, h->res, h->part, h->clk};
  for (void* q : ps)
    if (q) (void)hipFree(q);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return FVWB_OK;
}

static long long err_blocks(long long nc) { return (nc + ERR_THREADS - 1) / ERR_THREADS; }

// compute_max_speed (and, with a clock, evolve's dt) as three launches; the
// per-block partials use the handle's partial-sum buffer (3 doubles per block,
// at most err_blocks(nc) blocks)
static int speed_blocks(long long nc) {
  return (int)std::max(1LL, std::min<long long>((nc + SPEED_THREADS - 1) / SPEED_THREADS, SPEED_MAX_BLOCKS));
}
static void launch_speed(fvwb_dgq* h, const double* u, double* res, Clock* ck, double t_end) {
  const int nb = speed_blocks(h->nc);
  hipLaunchKernelGGL(dgq_speed_pairs_kernel, dim3(nb), dim3(SPEED_THREADS), 0, h->stream, h->g, u, h->part,
                     (const Clock*)ck, t_end);
  hipLaunchKernelGGL(dgq_speed_csmin_kernel, dim3(nb), dim3(SPEED_THREADS), 0, h->stream, h->g, u, h->part,
                     (const Clock*)ck, t_end);
  hipLaunchKernelGGL(dgq_speed_final_kernel, dim3(1), dim3(SPEED_THREADS), 0, h->stream, h->g, u,
                     (const double*)h->part, nb, res, ck, t_end);
}

extern "C" int fvwb_dgq_create(fvwb_dgq_handle* out, const fvwb_dgq_params* p) {
  if (!out || !p) return failq(nullptr, FVWB_ERR_ARG, "NULL argument");
  *out = nullptr;
  if (p->nx < 1 || p->nx != p->ny)
    return failq(nullptr, FVWB_ERR_ARG,
                 "nx == ny >= 1 required (the reference's residual indexes the x faces with the y bounds)");
  if ((long long)p->nx * p->ny > (1ll << 24)) return failq(nullptr, FVWB_ERR_ARG, "more than 2^24 cells");
  if (p->order < 2 || p->order > 3) return failq(nullptr, FVWB_ERR_ARG, "order (mx = my) must be 2 or 3");
  if (p->flux_type < 0 || p->flux_type > 3) return failq(nullptr, FVWB_ERR_ARG, "flux_type must be 0..3");
  if (p->source < 1 || p->source > 3) return failq(nullptr, FVWB_ERR_ARG, "source must be 1, 2 or 3");
  if (p->grad_phi_case < 1 || p->grad_phi_case > 2) return failq(nullptr, FVWB_ERR_ARG, "grad_phi_case must be 1 or 2");
  if (p->bc < 1 || p->bc > 3) return failq(nullptr, FVWB_ERR_ARG, "bc must be 1, 2 or 3");
  if (p->solver < 0 || p->solver > 3) return failq(nullptr, FVWB_ERR_ARG, "solver must be 0..3");
  if (!(p->gamma > 1.0) || !(p->boxlen_x > 0.0) || !(p->boxlen_y > 0.0) || !(p->cfl > 0.0))
    return failq(nullptr, FVWB_ERR_ARG, "gamma > 1, positive box lengths and cfl required");
  fvwb_dgq* h = new (std::nothrow) fvwb_dgq;
  if (!h) return failq(nullptr, FVWB_ERR_OOM, "host allocation failed");
  h->p = *p;
  const int o = p->order;
  h->T = make_tab(o);
  const double dx = p->boxlen_x / (double)p->nx;
  const int gll = h->T.gll;
  // evolve's gll_w_1 (:650-655) in default REAL: 1e-10 vanishes against gll(gll-1)
  const double gll_w_1 = o == 1 ? 1. : 1. / (double)((float)(gll * (gll - 1)) + 1e-10f);
  const double c1 = 1. / (double)(2 * 4 + 1), c2 = gll_w_1 / 2.;
  h->g = Geo{p->nx, p->ny, p->flux_type, p->source, p->grad_phi_case, p->bc, p->ninit, p->use_limiter,
             p->gamma, p->boxlen_x, p->boxlen_y, 1. / dx, dx, p->eps, p->cfl, c1 < c2 ? c1 : c2};
  h->nc = (long long)p->nx * p->ny;
  h->ns = 4 * h->nc * o * o;
  const size_t sb = sizeof(double) * (size_t)h->ns, cb = sizeof(double) * (size_t)(h->nc * o * o);
  const size_t trb = sizeof(double) * (size_t)(32LL * o * h->nc);
  const size_t fb = sizeof(double) * (size_t)(4LL * o * (p->nx + 1) * p->ny);
  const size_t gb = sizeof(double) * (size_t)(4LL * o * p->nx * (p->ny + 1));
  bool ok = hipStreamCreate(&h->stream) == hipSuccess && hipMalloc(&h->du, sb) == hipSuccess &&
            hipMalloc(&h->s_in, sb) == hipSuccess && hipMalloc(&h->s_out, sb) == hipSuccess &&
            hipMalloc(&h->x, cb) == hipSuccess && hipMalloc(&h->y, cb) == hipSuccess &&
            hipMalloc(&h->xi, cb) == hipSuccess && hipMalloc(&h->yi, cb) == hipSuccess &&
            hipMalloc(&h->tr, trb) == hipSuccess && hipMalloc(&h->F, fb) == hipSuccess &&
            hipMalloc(&h->G, gb) == hipSuccess && hipMalloc(&h->res, 16 * sizeof(double)) == hipSuccess &&
            hipMalloc(&h->part, sizeof(double) * 12 * (size_t)err_blocks(h->nc)) == hipSuccess &&
            hipMalloc(&h->clk, sizeof(Clock)) == hipSuccess;
  for (int k = 0; ok && k < 6; ++k) ok = hipMalloc(&h->w[k], sb) == hipSuccess;
  if (!ok) {
    fvwb_dgq_destroy(h);
    return failq(nullptr, FVWB_ERR_OOM, "device allocation failed");
  }
  *out = h;
  return FVWB_OK;
}

static dim3 cells_grid(const fvwb_dgq* h) { return dim3((unsigned)((h->nc + 255) / 256)); }
static dim3 cell_grid(const fvwb_dgq* h) { return dim3((unsigned)((h->nc + CELL_THREADS - 1) / CELL_THREADS)); }
static dim3 finish_grid(const fvwb_dgq* h) { return dim3((unsigned)std::min<long long>((h->ns + 255) / 256, 8192)); }

// compute_update of `in` into `out` with the stage combination e (nterms = 0:
// dudt itself); the clock gates and provides dt in the resident loop.
static int residual(fvwb_dgq* h, const double* in, const double* x, const double* y, const Epi& e, double* out,
                    const Clock* ck) {
  const long long nf = (long long)(h->p.nx + 1) * h->p.ny * h->p.order + (long long)h->p.nx * (h->p.ny + 1) * h->p.order;
  const dim3 fg((unsigned)((nf + 255) / 256));
  if (h->p.order == 2) {
    hipLaunchKernelGGL(dgq_trace_kernel<2>, cells_grid(h), dim3(256), 0, h->stream, h->g, h->T, in, h->tr, ck);
    hipLaunchKernelGGL(dgq_face_kernel<2>, fg, dim3(256), 0, h->stream, h->g, h->tr, h->F, h->G, ck);
    hipLaunchKernelGGL(dgq_cell_kernel<2>, cell_grid(h), dim3(CELL_THREADS), 0, h->stream, h->g, h->T, in, x, y,
                       h->F, h->G, e, out, ck);
  } else {
    hipLaunchKernelGGL(dgq_trace_kernel<3>, cells_grid(h), dim3(256), 0, h->stream, h->g, h->T, in, h->tr, ck);
    hipLaunchKernelGGL(dgq_face_kernel<3>, fg, dim3(256), 0, h->stream, h->g, h->tr, h->F, h->G, ck);
    hipLaunchKernelGGL(dgq_cell_kernel<3>, cell_grid(h), dim3(CELL_THREADS), 0, h->stream, h->g, h->T, in, x, y,
                       h->F, h->G, e, out, ck);
  }
  HIPQ(h, hipGetLastError());
  return FVWB_OK;
}

static int limit(fvwb_dgq* h, double* u, const Clock* ck) {
  if (!h->p.use_limiter) return FVWB_OK;
  if (h->p.order == 2)
    hipLaunchKernelGGL(dgq_limit_kernel<2>, cells_grid(h), dim3(256), 0, h->stream, h->g, h->T, u, ck);
  else
    hipLaunchKernelGGL(dgq_limit_kernel<3>, cells_grid(h), dim3(256), 0, h->stream, h->g, h->T, u, ck);
  HIPQ(h, hipGetLastError());
  return FVWB_OK;
}

static int project(fvwb_dgq* h, const double* src, double* dst, bool to_nodes) {
  if (h->p.order == 2) {
    if (to_nodes)
      hipLaunchKernelGGL((dgq_project_kernel<2, true>), cells_grid(h), dim3(256), 0, h->stream, h->g, h->T, src, dst);
    else
      hipLaunchKernelGGL((dgq_project_kernel<2, false>), cells_grid(h), dim3(256), 0, h->stream, h->g, h->T, src, dst);
  } else {
    if (to_nodes)
      hipLaunchKernelGGL((dgq_project_kernel<3, true>), cells_grid(h), dim3(256), 0, h->stream, h->g, h->T, src, dst);
    else
      hipLaunchKernelGGL((dgq_project_kernel<3, false>), cells_grid(h), dim3(256), 0, h->stream, h->g, h->T, src, dst);
  }
  HIPQ(h, hipGetLastError());
  return FVWB_OK;
}

static size_t state_bytes(const fvwb_dgq* h) { return sizeof(double) * (size_t)h->ns; }
static size_t coord_bytes(const fvwb_dgq* h) { return sizeof(double) * (size_t)(h->nc * h->p.order * h->p.order); }

extern "C" int fvwb_dgq_gl_quadrature(int32_t n, double* x_quad, double* w_quad) {
  if (!x_quad || !w_quad) return failq(nullptr, FVWB_ERR_ARG, "NULL argument");
  if (n < 1 || n > 6) return failq(nullptr, FVWB_ERR_ARG, "n must be 1..6 (legendre is defined up to 6)");
  h_gl_quadrature(n, x_quad, w_quad);
  return FVWB_OK;
}

extern "C" int fvwb_dgq_compute_update(fvwb_dgq_handle h, const double* delta_u, const double* x, const double* y,
                                       const double* u_eq, double* dudt) {
  (void)u_eq;      // unused by the reference (:1141)
  if (!h || !delta_u || !x || !y || !dudt) return failq(h, FVWB_ERR_ARG, "NULL argument");
  HIPQ(h, hipMemcpyAsync(h->s_in, delta_u, state_bytes(h), hipMemcpyHostToDevice, h->stream));
  HIPQ(h, hipMemcpyAsync(h->xi, x, coord_bytes(h), hipMemcpyHostToDevice, h->stream));
  HIPQ(h, hipMemcpyAsync(h->yi, y, coord_bytes(h), hipMemcpyHostToDevice, h->stream));
  Epi e{};
  int rc = residual(h, h->s_in, h->xi, h->yi, e, h->s_out, nullptr);
  if (rc) return rc;
  HIPQ(h, hipMemcpyAsync(dudt, h->s_out, state_bytes(h), hipMemcpyDeviceToHost, h->stream));
  HIPQ(h, hipStreamSynchronize(h->stream));
  return FVWB_OK;
}

extern "C" int fvwb_dgq_compute_max_speed(fvwb_dgq_handle h, const double* u, double* cs_max, double* v_xmax,
                                          double* v_ymax, double* speed_max) {
  if (!h || !u || !cs_max || !v_xmax || !v_ymax || !speed_max) return failq(h, FVWB_ERR_ARG, "NULL argument");
  HIPQ(h, hipMemcpyAsync(h->s_in, u, sizeof(double) * 4 * (size_t)h->nc, hipMemcpyHostToDevice, h->stream));
  launch_speed(h, h->s_in, h->res, nullptr, 0.0);
  HIPQ(h, hipGetLastError());
  double r[4];
  HIPQ(h, hipMemcpyAsync(r, h->res, sizeof r, hipMemcpyDeviceToHost, h->stream));
  HIPQ(h, hipStreamSynchronize(h->stream));
  *cs_max = r[0];
  *v_xmax = r[1];
  *v_ymax = r[2];
  *speed_max = r[3];
  return FVWB_OK;
}

extern "C" int fvwb_dgq_apply_limiter(fvwb_dgq_handle h, double* u) {
  if (!h || !u) return failq(h, FVWB_ERR_ARG, "NULL argument");
  HIPQ(h, hipMemcpyAsync(h->s_in, u, state_bytes(h), hipMemcpyHostToDevice, h->stream));
  int rc = limit(h, h->s_in, nullptr);
  if (rc) return rc;
  HIPQ(h, hipMemcpyAsync(u, h->s_in, state_bytes(h), hipMemcpyDeviceToHost, h->stream));
  HIPQ(h, hipStreamSynchronize(h->stream));
  return FVWB_OK;
}

static int project_host(fvwb_dgq* h, const double* src, double* dst, bool to_nodes) {
  if (!h || !src || !dst) return failq(h, FVWB_ERR_ARG, "NULL argument");
  HIPQ(h, hipMemcpyAsync(h->s_in, src, state_bytes(h), hipMemcpyHostToDevice, h->stream));
  int rc = project(h, h->s_in, h->s_out, to_nodes);
  if (rc) return rc;
  HIPQ(h, hipMemcpyAsync(dst, h->s_out, state_bytes(h), hipMemcpyDeviceToHost, h->stream));
  HIPQ(h, hipStreamSynchronize(h->stream));
  return FVWB_OK;
}

extern "C" int fvwb_dgq_modes_from_nodes(fvwb_dgq_handle h, const double* nodes, double* modes) {
  return project_host(h, nodes, modes, false);
}

extern "C" int fvwb_dgq_nodes_from_modes(fvwb_dgq_handle h, const double* modes, double* nodes) {
  return project_host(h, modes, nodes, true);
}

extern "C" int fvwb_dgq_upload(fvwb_dgq_handle h, const double* delta_u, const double* x, const double* y) {
  if (!h || !delta_u || !x || !y) return failq(h, FVWB_ERR_ARG, "NULL argument");
  const Clock zero{};
  HIPQ(h, hipMemcpyAsync(h->du, delta_u, state_bytes(h), hipMemcpyHostToDevice, h->stream));
  HIPQ(h, hipMemcpyAsync(h->x, x, coord_bytes(h), hipMemcpyHostToDevice, h->stream));
  HIPQ(h, hipMemcpyAsync(h->y, y, coord_bytes(h), hipMemcpyHostToDevice, h->stream));
  HIPQ(h, hipMemcpyAsync(h->clk, &zero, sizeof zero, hipMemcpyHostToDevice, h->stream));
  HIPQ(h, hipStreamSynchronize(h->stream));
  h->loaded = true;
  return FVWB_OK;
}

static int read_clock(fvwb_dgq* h, Clock* c) {
  HIPQ(h, hipMemcpyAsync(c, h->clk, sizeof *c, hipMemcpyDeviceToHost, h->stream));
  HIPQ(h, hipStreamSynchronize(h->stream));
  return FVWB_OK;
}

// the 'RK4' (:683-710) and 'SS4' (:711-735) coefficients, default-real literals
static const float kRK4[16] = {0.391752226571890f, 0.444370493651235f, 0.555629506348765f, 0.368410593050371f,
                               0.620101851488403f, 0.379898148511597f, 0.251891774271694f, 0.178079954393132f,
                               0.821920045606868f, 0.544974750228521f, 0.00683325884039f,  0.51723167208978f,
                               0.12759831133288f,  0.34833675773694f,  0.08460416338212f,  0.22600748319395f};
static const float kSS4[16] = {0.39175222700392f, 0.44437049406734f, 0.55562950593266f, 0.36841059262959f,
                               0.6201018513854f,  0.37989814861460f, 0.25189177424738f, 0.17807995410773f,
                               0.82192004589227f, 0.54497475021237f, 0.00683325884039f, 0.51723167208978f,
                               0.12759831133288f, 0.34833675773694f, 0.08460416338212f, 0.22600748319395f};

static Epi epi(std::initializer_list<std::pair<double, const double*>> terms, double cdt) {
  Epi e{};
  for (const auto& t : terms) {
    e.a[e.nterms] = t.first;
    e.X[e.nterms] = t.second;
    ++e.nterms;
  }
  e.cdt = cdt;
  return e;
}

// One step of evolve's loop (:666-769): dt, the configured solver with
// apply_limiter after its stages, t = t + dt.
static int step(fvwb_dgq* h, double t_end) {
  Clock* ck = h->clk;
  double* du = h->du;
  double *w1 = h->w[0], *w2 = h->w[1], *w3 = h->w[2], *w4 = h->w[3], *w5 = h->w[4], *l3 = h->w[5];
  launch_speed(h, du, nullptr, ck, t_end);
  HIPQ(h, hipGetLastError());
  int rc = 0;
  const double *x = h->x, *y = h->y;
  if (h->p.solver == FVWB_DGQ_SOLVER_EQL) {        // :672-681
    if ((rc = residual(h, du, x, y, epi({{1.0, du}}, 1.0), w1, ck)) || (rc = limit(h, w1, ck))) return rc;
    if ((rc = residual(h, w1, x, y, epi({{0.5, du}, {0.5, w1}}, 0.5), du, ck)) || (rc = limit(h, du, ck))) return rc;
  } else if (h->p.solver == FVWB_DGQ_SOLVER_DEB) { // :737-747: w1 = du + dt L(du), limited, becomes du
    if ((rc = residual(h, du, x, y, epi({{1.0, du}}, 1.0), du, ck)) || (rc = limit(h, du, ck))) return rc;
  } else {                                          // 'RK4' / 'SS4'
    const float* cf = h->p.solver == FVWB_DGQ_SOLVER_SS4 ? kSS4 : kRK4;
    double c[16];
    for (int k = 0; k < 16; ++k) c[k] = (double)cf[k];
    if ((rc = residual(h, du, x, y, epi({{1.0, du}}, c[0]), w1, ck)) || (rc = limit(h, w1, ck))) return rc;
    if ((rc = residual(h, w1, x, y, epi({{c[1], du}, {c[2], w1}}, c[3]), w2, ck)) || (rc = limit(h, w2, ck))) return rc;
    if ((rc = residual(h, w2, x, y, epi({{c[4], du}, {c[5], w2}}, c[6]), w3, ck)) || (rc = limit(h, w3, ck))) return rc;
    // w4 = c7 u + c8 w3 + c9 dt L(w3).  The reference evaluates L(w3) once more
    // (:700, :723) for w5 = c10 u + c11 w2 + c12 w3 + c13 w4 + c14 dt L(w3),
    // with w4 limited; here the first evaluation also keeps L(w3) and the sum
    // of w5's first three terms, and w5 is finished once w4 is limited.
    Epi e4 = epi({{c[7], du}, {c[8], w3}}, c[9]);
    e4.Y[0] = du; e4.Y[1] = w2; e4.Y[2] = w3;
    e4.b[0] = c[10]; e4.b[1] = c[11]; e4.b[2] = c[12];
    e4.npart = 3; e4.part = w5; e4.lval = l3;
    if ((rc = residual(h, w3, x, y, e4, w4, ck)) || (rc = limit(h, w4, ck))) return rc;
    hipLaunchKernelGGL(dgq_finish_kernel, finish_grid(h), dim3(256), 0, h->stream, h->ns, (const double*)w5, c[13],
                       (const double*)w4, c[14], (const double*)l3, w5, (const Clock*)ck);
    HIPQ(h, hipGetLastError());
    // then L(w4) for the new state
    if ((rc = residual(h, w4, x, y, epi({{1.0, w5}}, c[15]), du, ck)) || (rc = limit(h, du, ck))) return rc;
  }
  hipLaunchKernelGGL(dgq_tick_kernel, dim3(1), dim3(64), 0, h->stream, ck);
  HIPQ(h, hipGetLastError());
  return FVWB_OK;
}

static constexpr int32_t kGroupQ = 32;

extern "C" int fvwb_dgq_advance(fvwb_dgq_handle h, int32_t nsteps, double t_end) {
  if (!h) return failq(h, FVWB_ERR_ARG, "NULL handle");
  if (!h->loaded) return failq(h, FVWB_ERR_STATE, "advance before upload");
  if (nsteps < 0) return failq(h, FVWB_ERR_ARG, "nsteps must be >= 0");
  Clock c{};
  int rc = read_clock(h, &c);
  if (rc) return rc;
  int32_t done = 0;
  while (done < nsteps && !c.err && c.t < t_end) {
    const int32_t k = std::min(kGroupQ, nsteps - done);
    for (int32_t s = 0; s < k; ++s)
      if ((rc = step(h, t_end))) return rc;
    done += k;
    if ((rc = read_clock(h, &c))) return rc;
  }
  if (c.err) return failq(h, FVWB_ERR_NONPHYSICAL, "time step not finite and positive");
  return FVWB_OK;
}

extern "C" int fvwb_dgq_download(fvwb_dgq_handle h, double* delta_u) {
  if (!h || !delta_u) return failq(h, FVWB_ERR_ARG, "NULL argument");
  if (!h->loaded) return failq(h, FVWB_ERR_STATE, "download before upload");
  HIPQ(h, hipMemcpyAsync(delta_u, h->du, state_bytes(h), hipMemcpyDeviceToHost, h->stream));
  HIPQ(h, hipStreamSynchronize(h->stream));
  return FVWB_OK;
}

extern "C" int fvwb_dgq_get_time(fvwb_dgq_handle h, double* t, double* dt_last, int64_t* steps) {
  if (!h) return failq(h, FVWB_ERR_ARG, "NULL handle");
  Clock c{};
  if (h->loaded) {
    int rc = read_clock(h, &c);
    if (rc) return rc;
  }
  if (t) *t = c.t;
  if (dt_last) *dt_last = c.dt_last;
  if (steps) *steps = c.steps;
  return FVWB_OK;
}

extern "C" int fvwb_dgq_compute_error(fvwb_dgq_handle h, const double* nodes, const double* u_init, double* linf,
                                      double* l1, double* l2) {
  if (!h || !u_init || !linf || !l1 || !l2) return failq(h, FVWB_ERR_ARG, "NULL argument");
  if (!nodes && !h->loaded) return failq(h, FVWB_ERR_STATE, "no nodal values and no uploaded state");
  if (nodes) {
    HIPQ(h, hipMemcpyAsync(h->s_out, nodes, state_bytes(h), hipMemcpyHostToDevice, h->stream));
  } else {
    int rc = project(h, h->du, h->s_out, true);
    if (rc) return rc;
  }
  HIPQ(h, hipMemcpyAsync(h->s_in, u_init, state_bytes(h), hipMemcpyHostToDevice, h->stream));
  const long long nb = err_blocks(h->nc);
  if (h->p.order == 2)
    hipLaunchKernelGGL(dgq_error_kernel<2>, dim3((unsigned)nb), dim3(ERR_THREADS), 0, h->stream, h->g, h->T,
                       h->s_out, h->s_in, h->part);
  else
    hipLaunchKernelGGL(dgq_error_kernel<3>, dim3((unsigned)nb), dim3(ERR_THREADS), 0, h->stream, h->g, h->T,
                       h->s_out, h->s_in, h->part);
  hipLaunchKernelGGL(dgq_error_final_kernel, dim3(1), dim3(64), 0, h->stream, h->part, (int)nb, h->res);
  HIPQ(h, hipGetLastError());
  double r[12];
  HIPQ(h, hipMemcpyAsync(r, h->res, sizeof r, hipMemcpyDeviceToHost, h->stream));
  HIPQ(h, hipStreamSynchronize(h->stream));
  for (int v = 0; v < 4; ++v) {
    linf[v] = r[v];
    l1[v] = r[4 + v];
    l2[v] = r[8 + v];
  }
  return FVWB_OK;
}

// ---- compute_error without allocation ------------------------------------------
// The block partials of dgq_error_kernel live in this file-scope array instead
// of a handle's allocation.  Its host-side name is only a stand-in: the device
// address comes from hipGetSymbolAddress, once per device, and only that
// address is handed to kernels.  Every call on a GPU shares that GPU's array;
// each block overwrites its own 12 partials and the final pass reads only the
// blocks of the current call, so nothing needs clearing, but calls that
// overlap on one GPU race.
static __device__ __align__(256) double scratch[12 * (FVWB_DGQ_NOALLOC_MAX_CELLS / ERR_THREADS)];

namespace {
constexpr int kMaxDevices = 64;
std::atomic<double*> g_scratch_ptr[kMaxDevices];
std::mutex g_scratch_mu;

// the device address of `scratch` on the current device (nullptr on failure)
double* scratch_address(int dev) {
  double* ptr = g_scratch_ptr[dev].load(std::memory_order_acquire);
  if (ptr) return ptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  ptr = g_scratch_ptr[dev].load(std::memory_order_relaxed);
  if (ptr) return ptr;
  void* a = nullptr;
  if (hipGetSymbolAddress(&a, HIP_SYMBOL(scratch)) != hipSuccess || !a) return nullptr;
  ptr = static_cast<double*>(a);
  g_scratch_ptr[dev].store(ptr, std::memory_order_release);
  return ptr;
}

int fail_static(int code, const char* msg) {
  g_errq_static = msg;
  return code;
}
}  // namespace

extern "C" int fvwb_dgq_compute_error_noalloc(int32_t nx, int32_t ny, int32_t order, const double* nodes,
                                              const double* u_init, double* out, void* stream) {
  if (!nodes || !u_init || !out) return fail_static(FVWB_ERR_ARG, "NULL argument");
  if (nx < 1 || ny < 1) return fail_static(FVWB_ERR_ARG, "nx, ny >= 1 required");
  if (order < 2 || order > 3) return fail_static(FVWB_ERR_ARG, "order must be 2 or 3");
  const long long nc = (long long)nx * ny;
  if (nc > FVWB_DGQ_NOALLOC_MAX_CELLS)
    return fail_static(FVWB_ERR_ARG, "more cells than FVWB_DGQ_NOALLOC_MAX_CELLS (2^20): the scratch array "
                                     "holds the partial sums of that many cells");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return fail_static(FVWB_ERR_HIP, "hipGetDevice failed or device index out of range");
  double* part = scratch_address(dev);
  if (!part) return fail_static(FVWB_ERR_HIP, "hipGetSymbolAddress of the scratch array failed");
  hipStream_t st = static_cast<hipStream_t>(stream);
  Geo g{};
  g.nx = nx;
  g.ny = ny;
  const Tab T = make_tab(order);
  const long long nb = err_blocks(nc);
  if (order == 2)
    hipLaunchKernelGGL(dgq_error_kernel<2>, dim3((unsigned)nb), dim3(ERR_THREADS), 0, st, g, T, nodes, u_init, part);
  else
    hipLaunchKernelGGL(dgq_error_kernel<3>, dim3((unsigned)nb), dim3(ERR_THREADS), 0, st, g, T, nodes, u_init, part);
  hipLaunchKernelGGL(dgq_error_final_kernel, dim3(1), dim3(64), 0, st, (const double*)part, (int)nb, out);
  if (hipGetLastError() != hipSuccess) return fail_static(FVWB_ERR_HIP, "kernel launch failed");
  return FVWB_OK;
}
