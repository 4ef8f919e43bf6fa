// fvwb_dg1d.hip — the reference's 1D modal DG with source (dg_with_source.f90,
// legendre.f90, dg_commons.f90) on MI355X.
//
// dg1d_update_kernel is the drop-in for compute_update_exact_delta
// (dg_with_source.f90:1749-2031), the well-balanced deviation-form residual:
// one thread per cell, FP64, the reference's statement order with contraction
// off, so it returns the reference's dudt bit for bit.  What the reference
// evaluates with libm (the equilibrium exp at the faces, the quadrature and the
// basis values) is computed once on the host with the reference's expressions
// and handed to the kernel as tables; everything on the device is +, -, *, /
// and sqrt, which round identically on both sides.
#include <hip/hip_runtime.h>

#include <cmath>
#include <algorithm>
#include <cstring>
#include <new>
#include <string>

#include "../../include/fvwb.h"
#include "fvwb_checkpoint.h"
#include "fvwb_euler1d.h"

namespace dg1d {

constexpr int MAXN = 6;

// Basis and quadrature values of one set-up: phq[i*N+j] = phi_i(xi_j),
// dphq[i*N+j] = phi_i'(xi_j), wq[j], plr[2i] = phi_i(-1), plr[2i+1] = phi_i(+1).
struct Tables {
  double phq[MAXN * MAXN];
  double dphq[MAXN * MAXN];
  double wq[MAXN];
  double plr[2 * MAXN];
};

// ---- host: legendre.f90 as the reference evaluates it --------------------------
// x**k by binary powering, as the compiler's integer-power routine does.
static double ipow(double a, int b) {
  double r = 1;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return r;
}

// legendre.f90:1-24: phi_0..phi_2 normalised with FP32 square roots, P_3..P_6 plain
static double legendre(double x, int n) {
  x = std::fmin(std::fmax(x, -1.0), 1.0);
  switch (n) {
    case 0: return 1.0 * (double)std::sqrt(0.5f);
    case 1: return x * 0.5 * (double)std::sqrt(6.f);
    case 2: return 0.25 * (3.0 * ipow(x, 2) - 1.0) * (double)std::sqrt(10.f);
    case 3: return 0.5 * (5.0 * ipow(x, 3) - 3.0 * x);
    case 4: return 0.125 * (35.0 * ipow(x, 4) - 30.0 * ipow(x, 2) + 3.0);
    case 5: return 0.125 * (63.0 * ipow(x, 5) - 70.0 * ipow(x, 3) + 15.0 * x);
    case 6: return 1.0 / 16.0 * (231.0 * ipow(x, 6) - 315.0 * ipow(x, 4) + 105.0 * ipow(x, 2) - 5.0);
    default: return NAN;
  }
}

// legendre.f90:26-48
static double legendre_prime(double x, int n) {
  x = std::fmin(std::fmax(x, -1.0), 1.0);
  switch (n) {
    case 0: return 0.0;
    case 1: return (double)(1.0f * 0.5f * std::sqrt(6.f));
    case 2: return 6.0 * x * 0.25 * (double)std::sqrt(10.f);
    case 3: return 0.5 * (15.0 * ipow(x, 2) - 3.0);
    case 4: return 0.125 * (140.0 * ipow(x, 3) - 60.0 * x);
    case 5: return 0.125 * (315.0 * ipow(x, 4) - 210.0 * ipow(x, 2) + 15.0);
    case 6: return 1.0 / 16.0 * (1386.0 * ipow(x, 5) - 1260.0 * ipow(x, 3) + 210.0 * x);
    default: return NAN;
  }
}

// legendre.f90:77-127: FP32 literals up to three points, else Newton from the
// classical guess (50 steps), the weight 2(2n+1)/((1-x^2) P_n'^2) as written,
// and the second half mirrored from the first.
static void gl_quadrature(int n, double* x, double* w) {
  if (n == 1) {
    x[0] = 0.0;
    w[0] = 2.0;
    return;
  }
  if (n == 2) {
    x[0] = -(1. / (double)3 * (double)std::sqrt(3.f));
    x[1] = 1. / (double)3 * (double)std::sqrt(3.f);
    w[0] = w[1] = 1.;
    return;
  }
  if (n == 3) {
    x[0] = (double)(-(std::sqrt(3.f) / std::sqrt(5.f)));
    x[1] = 0.0;
    x[2] = (double)(std::sqrt(3.f) / std::sqrt(5.f));
    w[0] = w[2] = (double)(5.f / 9.f);
    w[1] = (double)(8.f / 9.f);
    return;
  }
  const double dpi = std::acos(-1.0);
  const float nf = (float)n;
  const float a = 1.0f - 0.125f / nf / nf + 0.125f / nf / nf / nf;
  for (int i = 1; i <= n; ++i) {
    double xx = (double)a * std::cos(dpi * (4.0 * (double)i - 1.0) / (4.0 * (double)n + 2.0));
    for (int it = 0; it < 50; ++it) xx = xx - legendre(xx, n) / legendre_prime(xx, n);
    x[i - 1] = xx;
    w[i - 1] = 2.0 * (2.0 * (double)n + 1.0) / (1.0 - ipow(x[i - 1], 2)) / ipow(legendre_prime(x[i - 1], n), 2);
  }
  for (int i = n / 2 + 1; i <= n; ++i) {
    x[i - 1] = -x[n - i];
    w[i - 1] = w[n - i];
  }
}

// ---- device ---------------------------------------------------------------------
// riemann_hllc (dg_with_source.f90:1318-1374) as written: wave speeds from the
// extreme velocities and the larger sound speed, acoustic star state, samples
// at x/t = 0 in the order left, left star, right star, right.
__device__ inline void riemann_hllc(double gamma, const double ul[3], const double ur[3], double f[3]) {
  double wl[3], wr[3];
  fv1d::primitive(gamma, ul, wl);
  fv1d::primitive(gamma, ur, wr);
  const double cl = sqrt(gamma * fmax(wl[2], 1e-10) / fmax(wl[0], 1e-10));
  const double cr = sqrt(gamma * fmax(wr[2], 1e-10) / fmax(wr[0], 1e-10));
  const double sl = fmin(wl[1], wr[1]) - fmax(cl, cr);
  const double sr = fmax(wl[1], wr[1]) + fmax(cl, cr);
  const double dl = wl[0] * (wl[1] - sl);
  const double dr = wr[0] * (sr - wr[1]);
  const double us = (dr * wr[1] + dl * wl[1] + (wl[2] - wr[2])) / (dl + dr);
  const double ps = (dr * wl[2] + dl * wr[2] + dl * dr * (wl[1] - wr[1])) / (dl + dr);
  double g0, g1, g2, e;
  if (sl > 0.0) {
    g0 = wl[0]; g1 = wl[1]; g2 = wl[2]; e = ul[2];
  } else if (us > 0.0) {
    g0 = wl[0] * (sl - wl[1]) / (sl - us);
    g1 = us; g2 = ps;
    e = ((sl - wl[1]) * ul[2] - wl[2] * wl[1] + ps * us) / (sl - us);
  } else if (sr > 0.0) {
    g0 = wr[0] * (sr - wr[1]) / (sr - us);
    g1 = us; g2 = ps;
    e = ((sr - wr[1]) * ur[2] - wr[2] * wr[1] + ps * us) / (sr - us);
  } else {
    g0 = wr[0]; g1 = wr[1]; g2 = wr[2]; e = ur[2];
  }
  f[0] = g0 * g1;
  f[1] = g0 * g1 * g1 + g2;
  f[2] = g1 * (e + g2);
}

__device__ inline void face_flux(double gamma, int riemann, const double ul[3], const double ur[3], double f[3]) {
  if (riemann == FVWB_DG1D_RIEMANN_HLLC) {
    riemann_hllc(gamma, ul, ur, f);
  } else {
    // riemann_llf (:1299-1316): compute_llflux with the fluxes taken inside
    double fl[3], fr[3];
    fv1d::flux(gamma, ul, fl);
    fv1d::flux(gamma, ur, fr);
    fv1d::llflux(gamma, ul, ur, fl, fr, f);
  }
}

// compute_source (:1196-1211): phi = x, g = 1
__device__ inline void source_term(double gamma, const double u[3], double s[3]) {
  double w[3];
  fv1d::primitive(gamma, u, w);
  s[0] = 0;
  s[1] = -w[0];
  s[2] = -w[0] * w[1];
}

struct Geometry {
  const double* ueq;   // equilibrium at the Gauss points, AoS(3,N,nx)
  const double* ufe;   // equilibrium conservative state at the nx+1 faces
  const double* ffe;   // its physical flux there
  int nx;
  int riemann, source;
  double gamma, oneoverdx;
};

// dudt of cell c (AoS(3,N) into out) for the modes du, AoS(3,N,nx).
template <int N>
__device__ inline void cell_residual(const double* du, const Geometry& G, const Tables& T, int c,
                                     double out[3 * N]) {
  const int nx = G.nx;
  const double gamma = G.gamma, oneoverdx = G.oneoverdx;
  const int riemann = G.riemann, source = G.source;
  const double* ueq = G.ueq;
  const double* ufe = G.ufe;
  const double* ffe = G.ffe;
  if (c == 0 || c == nx - 1) {   // dudt(:,:,1) = dudt(:,:,nx) = 0 (:2027-2028)
    for (int k = 0; k < 3 * N; ++k) out[k] = 0;
    return;
  }
  double m[3 * N];
  for (int k = 0; k < 3 * N; ++k) m[k] = du[(size_t)3 * N * c + k];

  // traces at xi = -1 and +1 (:1851-1872), summed from 0 mode by mode
  auto trace = [&](const double* a, int side, double t[3]) {
    t[0] = t[1] = t[2] = 0.;
    for (int i = 0; i < N; ++i)
      for (int v = 0; v < 3; ++v) t[v] = t[v] + a[3 * i + v] * T.plr[2 * i + side];
  };
  double F0[3], F1[3];
  {
    double a[3 * N], t[3], ua[3], ub[3];
    for (int k = 0; k < 3 * N; ++k) a[k] = du[(size_t)3 * N * (c - 1) + k];
    trace(a, 1, t);                                   // u_delta_right of cell c-1
    for (int v = 0; v < 3; ++v) ua[v] = ufe[3 * c + v] + t[v];
    trace(m, 0, t);                                   // u_delta_left of cell c
    for (int v = 0; v < 3; ++v) ub[v] = ufe[3 * c + v] + t[v];
    face_flux(gamma, riemann, ua, ub, F0);
    trace(m, 1, t);
    for (int v = 0; v < 3; ++v) ua[v] = ufe[3 * (c + 1) + v] + t[v];
    for (int k = 0; k < 3 * N; ++k) a[k] = du[(size_t)3 * N * (c + 1) + k];
    trace(a, 0, t);
    for (int v = 0; v < 3; ++v) ub[v] = ufe[3 * (c + 1) + v] + t[v];
    face_flux(gamma, riemann, ua, ub, F1);
  }

  // values at the Gauss points (:1802-1815, :1964-1988)
  double fq[N][3], fqe[N][3], sq[N][3], sqe[N][3];
  for (int j = 0; j < N; ++j) {
    double uq[3] = {0.0, 0.0, 0.0}, us[3];
    for (int i = 0; i < N; ++i)
      for (int v = 0; v < 3; ++v) uq[v] = uq[v] + m[3 * i + v] * T.phq[i * N + j];
    const double* e = ueq + (size_t)3 * (N * c + j);
    for (int v = 0; v < 3; ++v) us[v] = e[v] + uq[v];
    fv1d::flux(gamma, us, fq[j]);
    fv1d::flux(gamma, e, fqe[j]);
    if (source == 2) {
      source_term(gamma, us, sq[j]);
      source_term(gamma, e, sqe[j]);
    }
  }
  const double* fe0 = ffe + 3 * c;
  const double* fe1 = ffe + 3 * (c + 1);
  for (int i = 0; i < N; ++i) {
    double fv[3] = {0.0, 0.0, 0.0}, fve[3] = {0.0, 0.0, 0.0};
    double sv[3] = {0.0, 0.0, 0.0}, sve[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < N; ++j)
      for (int v = 0; v < 3; ++v) {
        fv[v] = fv[v] + fq[j][v] * T.dphq[i * N + j] * T.wq[j];
        fve[v] = fve[v] + fqe[j][v] * T.dphq[i * N + j] * T.wq[j];
        if (source == 2) {
          sv[v] = sv[v] + sq[j][v] * T.phq[i * N + j] * T.wq[j] * 0.5;
          sve[v] = sve[v] + sqe[j][v] * T.phq[i * N + j] * T.wq[j] * 0.5;
        }
      }
    const double pl = T.plr[2 * i], pr = T.plr[2 * i + 1];
    for (int v = 0; v < 3; ++v)   // :2015-2025
      out[3 * i + v] = oneoverdx * fv[v] - oneoverdx * fve[v] - oneoverdx * (F1[v] * pr - F0[v] * pl) +
                       oneoverdx * (fe1[v] * pr - fe0[v] * pl) + sv[v] - sve[v];
  }
}

// The drop-in: one thread per cell.
template <int N>
__global__ __launch_bounds__(256) void dg1d_update_kernel(const double* __restrict__ du, Geometry G, Tables T,
                                                          double* __restrict__ dudt) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= G.nx) return;
  double out[3 * N];
  cell_residual<N>(du, G, T, c, out);
  for (int k = 0; k < 3 * N; ++k) dudt[(size_t)3 * N * c + k] = out[k];
}

// ---- the main program's time loop, integrator 'RKi' ------------------------------
struct Clock {
  double t;
  double dt_last;
  int32_t steps;
  uint32_t err;   // bit 0: the wave-speed maximum or dt was not finite and positive
};

struct RunArgs {
  double* d;                 // modes, AoS(3,N,nx)
  double *w1, *w2, *w3, *w4, *acc;
  double* uinit;             // the nodal reconstruction uinit, AoS(3,N,nx)
  const double* phx;         // phi_k(x_quad(c,i)) at [(c*N + i)*N + k]
  Clock* clk;
  int32_t step_limit;        // stop after this many steps in total
  double t_end;
  double dx;
  double coef[15];           // 'RKi' literals rounded to FP32, in the order of the stages
};

// One workgroup runs the whole loop; its threads stride over the cells, and a
// workgroup barrier separates the stages (each stage reads its input's
// neighbours).  The stage arrays live in device memory, where one workgroup's
// writes are visible to its own threads after the barrier.
template <int N>
__global__ __launch_bounds__(256) void dg1d_run_kernel(Geometry G, Tables T, RunArgs R) {
  __shared__ double red[4];
  const int tid = threadIdx.x;
  const int nx = G.nx;
  const size_t M = (size_t)3 * N;
  double t = R.clk->t, dtl = R.clk->dt_last;
  int32_t steps = R.clk->steps;
  uint32_t err = R.clk->err;
  const double* k = R.coef;
  while (err == 0 && steps < R.step_limit && t < R.t_end) {
    // cmax over the cells of |v|+c at the first Gauss point (compute_max_speed :1136-1152)
    double cm = 0.0;
    for (int c = tid; c < nx; c += blockDim.x) cm = fmax(cm, fv1d::speed(G.gamma, R.uinit + M * c));
    for (int off = 32; off > 0; off >>= 1) cm = fmax(cm, __shfl_xor(cm, off));
    if ((tid & 63) == 0) red[tid >> 6] = cm;
    __syncthreads();
    double cmax = 0.0;
    for (int wv = 0; wv < (int)(blockDim.x >> 6); ++wv) cmax = fmax(cmax, red[wv]);
    __syncthreads();
    const double dt = (double)0.9f * R.dx / cmax / (2.0 * (double)N + 1.0);   // :197
    if (!(dt > 0.0) || !isfinite(dt)) {
      err = 1;
      break;
    }
    double L[3 * N];
    // :264-288, each line as written: (c*dt)*dudt added last
    for (int c = tid; c < nx; c += blockDim.x) {
      cell_residual<N>(R.d, G, T, c, L);
      for (int q = 0; q < 3 * N; ++q) R.w1[M * c + q] = R.d[M * c + q] + k[0] * dt * L[q];
    }
    __syncthreads();
    for (int c = tid; c < nx; c += blockDim.x) {
      cell_residual<N>(R.w1, G, T, c, L);
      for (int q = 0; q < 3 * N; ++q)
        R.w2[M * c + q] = k[1] * R.d[M * c + q] + k[2] * R.w1[M * c + q] + k[3] * dt * L[q];
    }
    __syncthreads();
    for (int c = tid; c < nx; c += blockDim.x) {
      cell_residual<N>(R.w2, G, T, c, L);
      for (int q = 0; q < 3 * N; ++q)
        R.w3[M * c + q] = k[4] * R.d[M * c + q] + k[5] * R.w2[M * c + q] + k[6] * dt * L[q];
    }
    __syncthreads();
    for (int c = tid; c < nx; c += blockDim.x) {
      cell_residual<N>(R.w3, G, T, c, L);
      for (int q = 0; q < 3 * N; ++q) {
        R.w4[M * c + q] = k[7] * R.d[M * c + q] + k[8] * R.w3[M * c + q] + k[9] * dt * L[q];
        R.acc[M * c + q] = k[10] * R.w2[M * c + q] + k[11] * R.w3[M * c + q] + k[12] * dt * L[q];
      }
    }
    __syncthreads();
    for (int c = tid; c < nx; c += blockDim.x) {
      cell_residual<N>(R.w4, G, T, c, L);
      for (int q = 0; q < 3 * N; ++q)
        R.d[M * c + q] = R.acc[M * c + q] + k[13] * R.w4[M * c + q] + k[14] * dt * L[q];
    }
    __syncthreads();
    // uinit = u_eq + the modes at legendre(x_quad) (:303-330), own cell only
    for (int c = tid; c < nx; c += blockDim.x)
      for (int i = 0; i < N; ++i)
        for (int v = 0; v < 3; ++v) {
          double a = 0.0;
          for (int q = 0; q < N; ++q) a = a + R.d[M * c + 3 * q + v] * R.phx[((size_t)c * N + i) * N + q];
          R.uinit[M * c + 3 * i + v] = G.ueq[M * c + 3 * i + v] + a;
        }
    __syncthreads();
    t = t + dt;
    dtl = dt;
    ++steps;
  }
  if (tid == 0) {
    R.clk->t = t;
    R.clk->dt_last = dtl;
    R.clk->steps = steps;
    R.clk->err = err;
  }
}

}  // namespace dg1d

using namespace dg1d;

struct fvwb_dg1d {
  fvwb_dg1d_params p{};
  hipStream_t stream = nullptr;
  double* d_du = nullptr;     // modes, AoS(3,n,nx)
  double* d_ueq = nullptr;    // equilibrium at the Gauss points, AoS(3,n,nx)
  double* d_dudt = nullptr;
  double* d_face = nullptr;   // ufe then ffe, 3*(nx+1) each
  double* d_run = nullptr;    // the time loop's modes, w1..w4, acc and uinit
  double* d_phx = nullptr;    // phi_k at the physical x of each Gauss point, (nx, n, n)
  Clock* d_clk = nullptr;
  bool loaded = false;
  Tables tab{};
  double coef[15];
  std::string err;
};

static thread_local std::string g_errdg1d;

static int faildg1d(fvwb_dg1d* h, int code, const std::string& msg) {
  (h ? h->err : g_errdg1d) = msg;
  return code;
}

#define HIPD(h, call)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) return faildg1d(h, FVWB_ERR_HIP, std::string(#call) + ": " + \
                                                           hipGetErrorString(e_));       \
  } while (0)

extern "C" int fvwb_dg1d_params_reference(fvwb_dg1d_params* p, int32_t nx) {
  if (!p) return faildg1d(nullptr, FVWB_ERR_ARG, "NULL params");
  std::memset(p, 0, sizeof *p);
  p->nvar = 3;
  p->nx = nx;
  p->n = 3;
  p->riemann = FVWB_DG1D_RIEMANN_HLLC;
  p->source = 2;
  p->gamma = (double)1.4f;
  p->boxlen = 1.0;
  return FVWB_OK;
}

extern "C" const char* fvwb_dg1d_last_error(fvwb_dg1d_handle h) { return h ? h->err.c_str() : g_errdg1d.c_str(); }

extern "C" int fvwb_dg1d_destroy(fvwb_dg1d_handle h) {
  if (!h) return FVWB_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  void* ps[] = {h->d_du, h->d_ueq, h->d_dudt, h->d_face, h->d_run, h->d_phx, h->d_clk};
  for (void* q : ps)
    if (q) (void)hipFree(q);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return FVWB_OK;
}

extern "C" int fvwb_dg1d_create(fvwb_dg1d_handle* out, const fvwb_dg1d_params* p) {
  if (!out || !p) return faildg1d(nullptr, FVWB_ERR_ARG, "NULL argument");
  *out = nullptr;
  if (p->nvar != 3) return faildg1d(nullptr, FVWB_ERR_ARG, "nvar must be 3");
  if (p->nx < 3) return faildg1d(nullptr, FVWB_ERR_ARG, "nx must be >= 3");
  if (p->n < 1 || p->n > MAXN) return faildg1d(nullptr, FVWB_ERR_ARG, "n must be 1..6");
  if (p->riemann != FVWB_DG1D_RIEMANN_LLF && p->riemann != FVWB_DG1D_RIEMANN_HLLC)
    return faildg1d(nullptr, FVWB_ERR_ARG, "riemann must be 1 (LLF) or 2 (HLLC)");
  if (p->source != 1 && p->source != 2) return faildg1d(nullptr, FVWB_ERR_ARG, "source must be 1 or 2");
  if (!(p->gamma > 1.0) || !(p->boxlen > 0.0)) return faildg1d(nullptr, FVWB_ERR_ARG, "gamma > 1, boxlen > 0");
  if ((size_t)p->nx * (size_t)p->n * 3 > ((size_t)1 << 31))
    return faildg1d(nullptr, FVWB_ERR_ARG, "nx*n too large");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return faildg1d(nullptr, FVWB_ERR_HIP, "no HIP device");
  fvwb_dg1d* h = new (std::nothrow) fvwb_dg1d;
  if (!h) return faildg1d(nullptr, FVWB_ERR_OOM, "host allocation failed");
  h->p = *p;
  const int n = p->n, nx = p->nx;
  double chsi[MAXN], w[MAXN];
  gl_quadrature(n, chsi, w);
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) {
      h->tab.phq[i * n + j] = legendre(chsi[j], i);
      h->tab.dphq[i * n + j] = legendre_prime(chsi[j], i);
    }
    h->tab.wq[i] = w[i];
    h->tab.plr[2 * i] = legendre(-1.0, i);
    h->tab.plr[2 * i + 1] = legendre(1.0, i);
  }
  // equilibrium at the faces (get_eq_solution :1031-1048 at x = (i-1)*dx), its flux
  const double dx = p->boxlen / (double)nx;
  const size_t nf = (size_t)nx + 1;
  double* face = new (std::nothrow) double[6 * nf];
  if (!face) {
    delete h;
    return faildg1d(nullptr, FVWB_ERR_OOM, "host allocation failed");
  }
  for (size_t i = 0; i < nf; ++i) {
    const double x = (double)i * dx;
    const double wf[3] = {std::exp(-x), 0, std::exp(-x)};
    fv1d::conservative(p->gamma, wf, &face[3 * i]);
    fv1d::flux(p->gamma, &face[3 * i], &face[3 * nf + 3 * i]);
  }
  // phi_k at x_quad = xcell + dx/2*chsi_i, the physical coordinate the main
  // program's reconstruction passes to legendre (:314-323)
  const size_t nphx = (size_t)nx * n * n;
  double* phx = new (std::nothrow) double[nphx];
  if (!phx) {
    delete[] face;
    delete h;
    return faildg1d(nullptr, FVWB_ERR_OOM, "host allocation failed");
  }
  for (int c = 0; c < nx; ++c) {
    const double xcell = ((double)(c + 1) - 0.5) * dx;
    for (int i = 0; i < n; ++i)
      for (int q = 0; q < n; ++q) phx[((size_t)c * n + i) * n + q] = legendre(xcell + dx / 2.0 * chsi[i], q);
  }
  static const float rki[15] = {0.391752226571890f, 0.444370493651235f, 0.555629506348765f, 0.368410593050371f,
                                0.620101851488403f, 0.379898148511597f, 0.251891774271694f, 0.178079954393132f,
                                0.821920045606868f, 0.544974750228521f, 0.517231671970585f, 0.096059710526147f,
                                0.063692468666290f, 0.386708617503269f, 0.226007483236906f};
  for (int i = 0; i < 15; ++i) h->coef[i] = (double)rki[i];
  const size_t bytes = sizeof(double) * 3 * (size_t)n * (size_t)nx;
  bool ok = hipStreamCreate(&h->stream) == hipSuccess && hipMalloc(&h->d_du, bytes) == hipSuccess &&
            hipMalloc(&h->d_ueq, bytes) == hipSuccess && hipMalloc(&h->d_dudt, bytes) == hipSuccess &&
            hipMalloc(&h->d_face, sizeof(double) * 6 * nf) == hipSuccess &&
            hipMalloc(&h->d_run, 7 * bytes) == hipSuccess && hipMalloc(&h->d_phx, sizeof(double) * nphx) == hipSuccess &&
            hipMalloc(&h->d_clk, sizeof(Clock)) == hipSuccess &&
            hipMemcpy(h->d_face, face, sizeof(double) * 6 * nf, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(h->d_phx, phx, sizeof(double) * nphx, hipMemcpyHostToDevice) == hipSuccess;
  delete[] face;
  delete[] phx;
  if (!ok) {
    fvwb_dg1d_destroy(h);
    return faildg1d(nullptr, FVWB_ERR_OOM, "device allocation failed");
  }
  *out = h;
  return FVWB_OK;
}

static Geometry geometry(const fvwb_dg1d* h, const double* ueq) {
  const fvwb_dg1d_params& p = h->p;
  const size_t nf = (size_t)p.nx + 1;
  Geometry G;
  G.ueq = ueq;
  G.ufe = h->d_face;
  G.ffe = h->d_face + 3 * nf;
  G.nx = p.nx;
  G.riemann = p.riemann;
  G.source = p.source;
  G.gamma = p.gamma;
  G.oneoverdx = 1. / (p.boxlen / (double)p.nx);
  return G;
}

template <int N>
static void launch_update(fvwb_dg1d* h) {
  hipLaunchKernelGGL(dg1d_update_kernel<N>, dim3((h->p.nx + 255) / 256), dim3(256), 0, h->stream, h->d_du,
                     geometry(h, h->d_ueq), h->tab, h->d_dudt);
}

extern "C" int fvwb_dg1d_compute_update_exact_delta(fvwb_dg1d_handle h, const double* delta_u, const double* u_eq,
                                                    double* dudt) {
  if (!h || !delta_u || !u_eq || !dudt) return faildg1d(h, FVWB_ERR_ARG, "NULL argument");
  const size_t bytes = sizeof(double) * 3 * (size_t)h->p.n * (size_t)h->p.nx;
  HIPD(h, hipMemcpyAsync(h->d_du, delta_u, bytes, hipMemcpyHostToDevice, h->stream));
  HIPD(h, hipMemcpyAsync(h->d_ueq, u_eq, bytes, hipMemcpyHostToDevice, h->stream));
  switch (h->p.n) {
    case 1: launch_update<1>(h); break;
    case 2: launch_update<2>(h); break;
    case 3: launch_update<3>(h); break;
    case 4: launch_update<4>(h); break;
    case 5: launch_update<5>(h); break;
    default: launch_update<6>(h); break;
  }
  HIPD(h, hipGetLastError());
  HIPD(h, hipMemcpyAsync(dudt, h->d_dudt, bytes, hipMemcpyDeviceToHost, h->stream));
  HIPD(h, hipStreamSynchronize(h->stream));
  return FVWB_OK;
}

// ---- the time loop ----------------------------------------------------------------
static double* run_array(fvwb_dg1d* h, int k) {
  return h->d_run + (size_t)k * 3 * (size_t)h->p.n * (size_t)h->p.nx;
}

extern "C" int fvwb_dg1d_upload(fvwb_dg1d_handle h, const double* delta_u, const double* u_eq, const double* u_init) {
  if (!h || !delta_u || !u_eq || !u_init) return faildg1d(h, FVWB_ERR_ARG, "NULL argument");
  const size_t bytes = sizeof(double) * 3 * (size_t)h->p.n * (size_t)h->p.nx;
  HIPD(h, hipMemcpyAsync(run_array(h, 0), delta_u, bytes, hipMemcpyHostToDevice, h->stream));
  HIPD(h, hipMemcpyAsync(h->d_ueq, u_eq, bytes, hipMemcpyHostToDevice, h->stream));
  HIPD(h, hipMemcpyAsync(run_array(h, 6), u_init, bytes, hipMemcpyHostToDevice, h->stream));
  const Clock zero{0.0, 0.0, 0, 0};
  HIPD(h, hipMemcpyAsync(h->d_clk, &zero, sizeof zero, hipMemcpyHostToDevice, h->stream));
  HIPD(h, hipStreamSynchronize(h->stream));
  h->loaded = true;
  return FVWB_OK;
}

template <int N>
static void launch_run(fvwb_dg1d* h, int32_t limit, double t_end) {
  RunArgs R;
  R.d = run_array(h, 0);
  R.w1 = run_array(h, 1);
  R.w2 = run_array(h, 2);
  R.w3 = run_array(h, 3);
  R.w4 = run_array(h, 4);
  R.acc = run_array(h, 5);
  R.uinit = run_array(h, 6);
  R.phx = h->d_phx;
  R.clk = h->d_clk;
  R.step_limit = limit;
  R.t_end = t_end;
  R.dx = h->p.boxlen / (double)h->p.nx;
  for (int i = 0; i < 15; ++i) R.coef[i] = h->coef[i];
  hipLaunchKernelGGL(dg1d_run_kernel<N>, dim3(1), dim3(256), 0, h->stream, geometry(h, h->d_ueq), h->tab, R);
}

// Steps are launched in chunks of at most this many, so one launch stays short.
static constexpr int32_t kChunk = 2048;

extern "C" int fvwb_dg1d_advance(fvwb_dg1d_handle h, int32_t nsteps, double t_end) {
  if (!h) return faildg1d(h, FVWB_ERR_ARG, "NULL handle");
  if (!h->loaded) return faildg1d(h, FVWB_ERR_STATE, "advance before upload");
  if (nsteps < 0) return faildg1d(h, FVWB_ERR_ARG, "nsteps must be >= 0");
  Clock clk;
  HIPD(h, hipMemcpy(&clk, h->d_clk, sizeof clk, hipMemcpyDeviceToHost));
  const int64_t target = (int64_t)clk.steps + nsteps;
  for (;;) {
    if (clk.err) return faildg1d(h, FVWB_ERR_NONPHYSICAL, "wave-speed maximum or dt not finite and positive");
    if (!(clk.t < t_end) || clk.steps >= target) return FVWB_OK;
    const int32_t limit = (int32_t)std::min<int64_t>(target, (int64_t)clk.steps + kChunk);
    switch (h->p.n) {
      case 1: launch_run<1>(h, limit, t_end); break;
      case 2: launch_run<2>(h, limit, t_end); break;
      case 3: launch_run<3>(h, limit, t_end); break;
      case 4: launch_run<4>(h, limit, t_end); break;
      case 5: launch_run<5>(h, limit, t_end); break;
      default: launch_run<6>(h, limit, t_end); break;
    }
    HIPD(h, hipGetLastError());
    HIPD(h, hipMemcpyAsync(&clk, h->d_clk, sizeof clk, hipMemcpyDeviceToHost, h->stream));
    HIPD(h, hipStreamSynchronize(h->stream));
  }
}

extern "C" int fvwb_dg1d_download(fvwb_dg1d_handle h, double* delta_u, double* u_init) {
  if (!h || !delta_u) return faildg1d(h, FVWB_ERR_ARG, "NULL argument");
  if (!h->loaded) return faildg1d(h, FVWB_ERR_STATE, "download before upload");
  const size_t bytes = sizeof(double) * 3 * (size_t)h->p.n * (size_t)h->p.nx;
  HIPD(h, hipMemcpyAsync(delta_u, run_array(h, 0), bytes, hipMemcpyDeviceToHost, h->stream));
  if (u_init) HIPD(h, hipMemcpyAsync(u_init, run_array(h, 6), bytes, hipMemcpyDeviceToHost, h->stream));
  HIPD(h, hipStreamSynchronize(h->stream));
  return FVWB_OK;
}

extern "C" int fvwb_dg1d_get_time(fvwb_dg1d_handle h, double* t, double* dt_last, int32_t* steps) {
  if (!h) return faildg1d(h, FVWB_ERR_ARG, "NULL handle");
  Clock clk;
  HIPD(h, hipMemcpy(&clk, h->d_clk, sizeof clk, hipMemcpyDeviceToHost));
  if (t) *t = clk.t;
  if (dt_last) *dt_last = clk.dt_last;
  if (steps) *steps = clk.steps;
  return FVWB_OK;
}

// ---- checkpoint / resume: the modes, u_init, the equilibrium at the Gauss points, the clock
static int dg1d_segments(fvwb_dg1d* h, fvwb_ckpt::Segment seg[4]) {
  const size_t bytes = sizeof(double) * 3 * (size_t)h->p.n * (size_t)h->p.nx;
  seg[0] = {run_array(h, 0), bytes};
  seg[1] = {run_array(h, 6), bytes};
  seg[2] = {h->d_ueq, bytes};
  seg[3] = {h->d_clk, sizeof(Clock)};
  return 4;
}

extern "C" int fvwb_dg1d_checkpoint_save(fvwb_dg1d_handle h, const char* path) {
  if (!h || !path) return faildg1d(h, FVWB_ERR_ARG, "NULL argument");
  if (!h->loaded) return faildg1d(h, FVWB_ERR_STATE, "checkpoint before upload");
  fvwb_ckpt::Segment seg[4];
  const int n = dg1d_segments(h, seg);
  std::string err;
  const int rc = fvwb_ckpt::save(path, fvwb_ckpt::DG1D, &h->p, sizeof h->p, 0, seg, n, h->stream, err);
  if (rc) h->err = err;
  return rc;
}

extern "C" int fvwb_dg1d_checkpoint_load(fvwb_dg1d_handle h, const char* path) {
  if (!h || !path) return faildg1d(h, FVWB_ERR_ARG, "NULL argument");
  fvwb_ckpt::Segment seg[4];
  const int n = dg1d_segments(h, seg);
  std::string err;
  int32_t parity = 0;
  const int rc = fvwb_ckpt::load(path, fvwb_ckpt::DG1D, &h->p, sizeof h->p, &parity, seg, n, h->stream, err);
  h->loaded = rc == FVWB_OK;
  if (rc) h->err = err;
  return rc;
}
