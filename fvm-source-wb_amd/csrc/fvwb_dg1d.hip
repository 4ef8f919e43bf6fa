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
#include <cstring>
#include <new>
#include <string>

#include "../../include/fvwb.h"
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

// One thread per cell c.  du, ueq, dudt: AoS(3,N,nx); ufe, ffe: the
// equilibrium's conservative state and physical flux at the nx+1 faces.
template <int N>
__global__ __launch_bounds__(256) void dg1d_update_kernel(const double* __restrict__ du,
                                                          const double* __restrict__ ueq,
                                                          const double* __restrict__ ufe,
                                                          const double* __restrict__ ffe,
                                                          double* __restrict__ dudt, int nx, double gamma,
                                                          double oneoverdx, int riemann, int source, Tables T) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nx) return;
  double* out = dudt + (size_t)3 * N * c;
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

}  // namespace dg1d

using namespace dg1d;

struct fvwb_dg1d {
  fvwb_dg1d_params p{};
  hipStream_t stream = nullptr;
  double* d_du = nullptr;     // modes, AoS(3,n,nx)
  double* d_ueq = nullptr;    // equilibrium at the Gauss points, AoS(3,n,nx)
  double* d_dudt = nullptr;
  double* d_face = nullptr;   // ufe then ffe, 3*(nx+1) each
  Tables tab{};
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
  void* ps[] = {h->d_du, h->d_ueq, h->d_dudt, h->d_face};
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
  const size_t bytes = sizeof(double) * 3 * (size_t)n * (size_t)nx;
  bool ok = hipStreamCreate(&h->stream) == hipSuccess && hipMalloc(&h->d_du, bytes) == hipSuccess &&
            hipMalloc(&h->d_ueq, bytes) == hipSuccess && hipMalloc(&h->d_dudt, bytes) == hipSuccess &&
            hipMalloc(&h->d_face, sizeof(double) * 6 * nf) == hipSuccess &&
            hipMemcpy(h->d_face, face, sizeof(double) * 6 * nf, hipMemcpyHostToDevice) == hipSuccess;
  delete[] face;
  if (!ok) {
    fvwb_dg1d_destroy(h);
    return faildg1d(nullptr, FVWB_ERR_OOM, "device allocation failed");
  }
  *out = h;
  return FVWB_OK;
}

template <int N>
static void launch_update(fvwb_dg1d* h) {
  const fvwb_dg1d_params& p = h->p;
  const double dx = p.boxlen / (double)p.nx;
  const size_t nf = (size_t)p.nx + 1;
  hipLaunchKernelGGL(dg1d_update_kernel<N>, dim3((p.nx + 255) / 256), dim3(256), 0, h->stream, h->d_du, h->d_ueq,
                     h->d_face, h->d_face + 3 * nf, h->d_dudt, p.nx, p.gamma, 1. / dx, p.riemann, p.source, h->tab);
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
