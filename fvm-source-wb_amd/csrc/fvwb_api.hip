// fvwb_api.hip — C ABI of the single-instance 2D solver (include/fvwb.h).
//
// Replaces the external procedures compute_update_exact / compute_max_speed
// and the stepping loop of evolve (benchmark_2d.f90:221-279, 465-618).  The
// handle keeps the deviation dU = U - U_eq per variable (FP64 or FP32), the
// caller's primitive equilibrium (for download), separable FP64 equilibrium
// tables, a device step-control block and one HIP stream.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "../../include/fvwb.h"
#include "fvwb_checkpoint.h"
#include "fvwb_internal.h"

namespace fvwb {
hipError_t launch_fv2d_stage_f64(const StageArgs<double>& a, int recon, int flux, int bc, int mode, hipStream_t s);
hipError_t launch_fv2d_stage_f32(const StageArgs<float>& a, int recon, int flux, int bc, int mode, int cpl,
                                 hipStream_t s);
template <typename T>
hipError_t launch_aos_to_dev(const double* u, const double* w, T* out, int nx, int ny, int pitch, long long plane,
                             double gamma, hipStream_t s);
template <typename T>
hipError_t launch_dev_to_aos(const T* in, const double* w, double* u, int nx, int ny, int pitch, long long plane,
                             double gamma, bool add_eq, hipStream_t s);
hipError_t launch_max_speed_aos(const double* u, long long n, double gamma, unsigned long long* out, hipStream_t s);
hipError_t launch_plain_update(const double* u, double* dudt, int nx, int ny, double gamma, double boxlen_x,
                               double boxlen_y, hipStream_t s);
template <typename T>
hipError_t launch_cfl_init(const T* in, const double* rxc, const double* eyc, Ctrl* ctrl, int nx, int ny, int pitch,
                           long long plane, double gamma, double p_over_rho, double inv_gm1, hipStream_t s);
template <typename T>
hipError_t launch_init_gaussian(T* out, double* w, int nx, int ny, int pitch, long long plane, double dx, double dy,
                                double rho0, double p0, double kk, double gamma, double amp, double x0, double y0,
                                double width, hipStream_t s);
}  // namespace fvwb

using namespace fvwb;

struct fvwb_ctx {
  fvwb_params p{};
  hipStream_t stream = nullptr;
  // The four variables' rows of one grid row lie side by side: element
  // (v, i, j) of a state copy is at v plane + j pitch + i, with plane = nx
  // rounded up to 64 and pitch = 4 plane, so a strip's four row loads fall in
  // one span instead of a quarter of the copy apart.
  int pitch = 0;            // element stride between grid rows
  long long plane = 0;      // element stride between the variables
  size_t elems = 0;         // elements per state copy (pitch * ny)
  size_t esize = 8;
  void* buf[3] = {nullptr, nullptr, nullptr};  // A = dU^n, B, C
  double* w_eq = nullptr;                      // AoS(4,nx,ny) primitive equilibrium (device)
  double* stage_u = nullptr;                   // AoS staging of host arrays
  double* stage_w = nullptr;
  double* res_in = nullptr;                    // FP64 planes for the per-stage drop-in
  double* res_out = nullptr;
  double* rxf = nullptr;
  double* rxc = nullptr;
  double* eyc = nullptr;
  double* eyf = nullptr;
  double* eyf_ref = nullptr;   // y-face table at y = j*dx (the reference's face positions)
  Ctrl* ctrl = nullptr;
  unsigned long long* red = nullptr;
  int parity = 0;
  bool uploaded = false;
  Tune tune{16, 2};
  std::string err;
};

static thread_local std::string g_create_err;

#define FVWB_HIP(h, expr)                                                                        \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess) {                                                                      \
      (h)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                             \
      return e_ == hipErrorOutOfMemory ? FVWB_ERR_OOM : FVWB_ERR_HIP;                            \
    }                                                                                            \
  } while (0)

static int fail(fvwb_ctx* h, int code, const char* msg) {
  if (h) h->err = msg; else g_create_err = msg;
  return code;
}

static int validate(const fvwb_params* p, std::string& why) {
  if (!p) { why = "params is NULL"; return FVWB_ERR_ARG; }
  if (p->nvar != 4) { why = "nvar must be 4"; return FVWB_ERR_ARG; }
  if (p->nx < 4 || p->ny < 4) { why = "nx and ny must be >= 4"; return FVWB_ERR_ARG; }
  if ((long long)p->nx * p->ny > (1ll << 31)) { why = "nx*ny exceeds 2^31 cells"; return FVWB_ERR_ARG; }
  if (p->bc != FVWB_BC_FROZEN_RING && p->bc != FVWB_BC_REFLECT) { why = "unknown bc"; return FVWB_ERR_ARG; }
  if (p->recon != FVWB_RECON_FIRST_ORDER && p->recon != FVWB_RECON_MUSCL_MINMOD) { why = "unknown recon"; return FVWB_ERR_ARG; }
  if (p->flux != FVWB_FLUX_RUSANOV && p->flux != FVWB_FLUX_HLL) {
    why = "flux must be FVWB_FLUX_RUSANOV or FVWB_FLUX_HLL (HLLC is a DG flux)";
    return FVWB_ERR_ARG;
  }
  if (p->integrator != FVWB_RK2_HEUN && p->integrator != FVWB_SSPRK3) { why = "unknown integrator"; return FVWB_ERR_ARG; }
  if (p->precision != FVWB_FP64 && p->precision != FVWB_FP32) { why = "unknown precision"; return FVWB_ERR_ARG; }
  if (!(p->gamma > 1.0) || !(p->boxlen_x > 0) || !(p->boxlen_y > 0) || !(p->cfl > 0) || !(p->rho0 > 0) ||
      !(p->p0 > 0) || !(p->g >= 0)) { why = "non-physical constants"; return FVWB_ERR_ARG; }
  return FVWB_OK;
}

extern "C" int fvwb_params_reference_2d(fvwb_params* p, int32_t nx, int32_t ny, int32_t nequilibrium) {
  if (!p || nequilibrium < 1 || nequilibrium > 3) return FVWB_ERR_ARG;
  std::memset(p, 0, sizeof *p);
  p->nvar = 4; p->nx = nx; p->ny = ny;
  p->bc = FVWB_BC_FROZEN_RING; p->recon = FVWB_RECON_FIRST_ORDER; p->flux = FVWB_FLUX_RUSANOV;
  p->integrator = FVWB_RK2_HEUN; p->precision = FVWB_FP64;
  p->gamma = (double)1.4f;              // parameters_2d.f90:19
  p->boxlen_x = 1.0; p->boxlen_y = 1.0; // parameters_2d.f90:17-18
  p->cfl = 0.5;                         // parameters_2d.f90:20
  p->rho0 = nequilibrium == 1 ? 1.0 : (double)1.21f;  // benchmark_2d.f90:190-202
  p->p0 = 1.0; p->g = 1.0;
  return FVWB_OK;
}

static int alloc_dev(fvwb_ctx* h, void** ptr, size_t bytes) {
  FVWB_HIP(h, hipMalloc(ptr, bytes));
  return FVWB_OK;
}

extern "C" int fvwb_create(fvwb_handle* out, const fvwb_params* p) {
  if (!out) return fail(nullptr, FVWB_ERR_ARG, "handle pointer is NULL");
  *out = nullptr;
  std::string why;
  if (validate(p, why) != FVWB_OK) return fail(nullptr, FVWB_ERR_ARG, why.c_str());
  fvwb_ctx* h = new (std::nothrow) fvwb_ctx();
  if (!h) return fail(nullptr, FVWB_ERR_OOM, "host allocation failed");
  h->p = *p;
  const int nx = p->nx, ny = p->ny;
  h->plane = (nx + 63) / 64 * 64;
  h->pitch = 4 * (int)h->plane;
  h->elems = (size_t)h->pitch * ny;
  h->esize = p->precision == FVWB_FP64 ? 8 : 4;
  int rc = FVWB_OK;
  auto check = [&](int r) { if (r != FVWB_OK && rc == FVWB_OK) rc = r; };
  hipError_t e = hipStreamCreate(&h->stream);
  if (e != hipSuccess) { g_create_err = hipGetErrorString(e); delete h; return FVWB_ERR_HIP; }
  const size_t nc = (size_t)nx * ny;
  for (int b = 0; b < 3; ++b) check(alloc_dev(h, &h->buf[b], h->elems * h->esize));
  check(alloc_dev(h, (void**)&h->w_eq, 4 * nc * sizeof(double)));
  check(alloc_dev(h, (void**)&h->stage_u, 4 * nc * sizeof(double)));
  check(alloc_dev(h, (void**)&h->rxf, (nx + 1) * sizeof(double)));
  check(alloc_dev(h, (void**)&h->rxc, nx * sizeof(double)));
  check(alloc_dev(h, (void**)&h->eyc, ny * sizeof(double)));
  check(alloc_dev(h, (void**)&h->eyf, (ny + 1) * sizeof(double)));
  check(alloc_dev(h, (void**)&h->eyf_ref, (ny + 1) * sizeof(double)));
  check(alloc_dev(h, (void**)&h->ctrl, sizeof(Ctrl)));
  check(alloc_dev(h, (void**)&h->red, sizeof(unsigned long long)));
  if (rc != FVWB_OK) {
    g_create_err = h->err;
    fvwb_destroy(h);
    return rc;
  }
  // separable FP64 equilibrium tables: rho_eq(x,y) = rho0 e^{-kx} e^{-ky}
  const double dx = p->boxlen_x / nx, dy = p->boxlen_y / ny, k = p->rho0 * p->g / p->p0;
  std::string tmp;
  double* hx = new double[2 * (size_t)nx + 1];
  double* hy = new double[3 * (size_t)ny + 2];
  for (int i = 0; i <= nx; ++i) hx[i] = p->rho0 * std::exp(-(k * ((double)i * dx)));
  for (int i = 0; i < nx; ++i) hx[nx + 1 + i] = p->rho0 * std::exp(-(k * (((double)i + 0.5) * dx)));
  for (int j = 0; j < ny; ++j) hy[j] = std::exp(-(k * (((double)j + 0.5) * dy)));
  for (int j = 0; j <= ny; ++j) hy[ny + j] = std::exp(-(k * ((double)j * dy)));
  // benchmark_2d.f90:513 places the y faces at (j-1)*dx, not (j-1)*dy
  for (int j = 0; j <= ny; ++j) hy[2 * ny + 1 + j] = std::exp(-(k * ((double)j * dx)));
  hipError_t e1 = hipMemcpy(h->rxf, hx, (nx + 1) * sizeof(double), hipMemcpyHostToDevice);
  hipError_t e2 = hipMemcpy(h->rxc, hx + nx + 1, nx * sizeof(double), hipMemcpyHostToDevice);
  hipError_t e3 = hipMemcpy(h->eyc, hy, ny * sizeof(double), hipMemcpyHostToDevice);
  hipError_t e4 = hipMemcpy(h->eyf, hy + ny, (ny + 1) * sizeof(double), hipMemcpyHostToDevice);
  if (e4 == hipSuccess)
    e4 = hipMemcpy(h->eyf_ref, hy + 2 * ny + 1, (ny + 1) * sizeof(double), hipMemcpyHostToDevice);
  delete[] hx;
  delete[] hy;
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess || e4 != hipSuccess) {
    g_create_err = "copying equilibrium tables failed";
    fvwb_destroy(h);
    return FVWB_ERR_HIP;
  }
  // launch shape: rows per wave strip so that a 2048^2 grid yields >= 4k waves
  long long cells = (long long)nx * ny;
  // launch shape: short strips keep >= ~16 waves per CU on mid-size grids;
  // FP32 lanes own 2 columns (8-byte loads leave registers for a deeper ring)
  h->tune.rows = cells >= (1ll << 24) ? 16 : (cells >= (1ll << 20) ? 8 : 4);
  h->tune.cpl = 2;
  *out = h;
  return FVWB_OK;
}

extern "C" int fvwb_destroy(fvwb_handle h) {
  if (!h) return FVWB_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (int b = 0; b < 3; ++b) if (h->buf[b]) (void)hipFree(h->buf[b]);
  double* ds[] = {h->w_eq, h->stage_u, h->stage_w, h->res_in, h->res_out, h->rxf, h->rxc, h->eyc, h->eyf,
                   h->eyf_ref};
  for (double* d : ds) if (d) (void)hipFree(d);
  if (h->ctrl) (void)hipFree(h->ctrl);
  if (h->red) (void)hipFree(h->red);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return FVWB_OK;
}

extern "C" const char* fvwb_last_error(fvwb_handle h) {
  return h ? h->err.c_str() : g_create_err.c_str();
}

template <typename T>
static StageArgs<T> base_args(fvwb_ctx* h) {
  const fvwb_params& p = h->p;
  StageArgs<T> a{};
  a.rxf = h->rxf; a.rxc = h->rxc; a.eyc = h->eyc; a.eyf = h->eyf;
  a.ctrl = h->ctrl;
  a.plane = h->plane; a.pitch = h->pitch; a.nx = p.nx; a.ny = p.ny;
  a.rows = h->tune.rows;
  const int cpl = sizeof(T) == 8 ? 2 : h->tune.cpl;
  const int cols = 62 * cpl;   // columns per wave strip
  a.ntx = (p.nx + cols - 1) / cols;
  a.ntiles = a.ntx * ((p.ny + a.rows - 1) / a.rows);
  const double dx = p.boxlen_x / p.nx, dy = p.boxlen_y / p.ny;
  a.h = dx < dy ? dx : dy;
  a.cfl = p.cfl;
  a.gamma = p.gamma; a.g = p.g; a.p_over_rho = p.p0 / p.rho0; a.inv_gm1 = 1.0 / (p.gamma - 1.0);
  a.odx = 1.0 / dx; a.ody = 1.0 / dy;
  a.t_end = INFINITY;
  a.alpha = 0.0; a.beta = 1.0;
  return a;
}

static int ensure_residual_buffers(fvwb_ctx* h) {
  const size_t nc = (size_t)h->p.nx * h->p.ny;
  if (!h->stage_w) FVWB_HIP(h, hipMalloc((void**)&h->stage_w, 4 * nc * sizeof(double)));
  if (!h->res_in) FVWB_HIP(h, hipMalloc((void**)&h->res_in, h->elems * sizeof(double)));
  if (!h->res_out) FVWB_HIP(h, hipMalloc((void**)&h->res_out, h->elems * sizeof(double)));
  return FVWB_OK;
}

// Residual of the handle's scheme on host arrays; reference_faces selects the
// reference's y-face positions (j*dx) for the face equilibrium.
static int residual_host(fvwb_handle h, const double* u, const double* w_eq, double* dudt, bool reference_faces) {
  if (!h || !u || !w_eq || !dudt) return fail(h, FVWB_ERR_ARG, "NULL argument");
  int rc = ensure_residual_buffers(h);
  if (rc) return rc;
  const fvwb_params& p = h->p;
  const size_t bytes = 4 * (size_t)p.nx * p.ny * sizeof(double);
  FVWB_HIP(h, hipMemcpyAsync(h->stage_u, u, bytes, hipMemcpyHostToDevice, h->stream));
  FVWB_HIP(h, hipMemcpyAsync(h->stage_w, w_eq, bytes, hipMemcpyHostToDevice, h->stream));
  FVWB_HIP(h, launch_aos_to_dev<double>(h->stage_u, h->stage_w, h->res_in, p.nx, p.ny, h->pitch, h->plane, p.gamma,
                                        h->stream));
  StageArgs<double> a = base_args<double>(h);
  a.in = h->res_in; a.un = nullptr; a.out = h->res_out;
  if (reference_faces) a.eyf = h->eyf_ref;
  FVWB_HIP(h, launch_fv2d_stage_f64(a, p.recon, p.flux, p.bc, 0, h->stream));
  FVWB_HIP(h, launch_dev_to_aos<double>(h->res_out, nullptr, h->stage_u, p.nx, p.ny, h->pitch, h->plane, p.gamma,
                                        false, h->stream));
  FVWB_HIP(h, hipMemcpyAsync(dudt, h->stage_u, bytes, hipMemcpyDeviceToHost, h->stream));
  FVWB_HIP(h, hipStreamSynchronize(h->stream));
  return FVWB_OK;
}

extern "C" int fvwb_compute_update_exact(fvwb_handle h, const double* u, const double* w_eq, double* dudt) {
  return residual_host(h, u, w_eq, dudt, true);
}

extern "C" int fvwb_compute_residual(fvwb_handle h, const double* u, const double* w_eq, double* dudt) {
  return residual_host(h, u, w_eq, dudt, false);
}

// compute_update (benchmark_2d.f90:370-463), the plain non-well-balanced
// residual; w_eq is part of the reference's signature and unused there too.
extern "C" int fvwb_compute_update(fvwb_handle h, const double* u, const double* w_eq, double* dudt) {
  (void)w_eq;
  if (!h || !u || !dudt) return fail(h, FVWB_ERR_ARG, "NULL argument");
  int rc = ensure_residual_buffers(h);
  if (rc) return rc;
  const fvwb_params& p = h->p;
  const size_t bytes = 4 * (size_t)p.nx * p.ny * sizeof(double);
  FVWB_HIP(h, hipMemcpyAsync(h->stage_u, u, bytes, hipMemcpyHostToDevice, h->stream));
  FVWB_HIP(h, launch_plain_update(h->stage_u, h->stage_w, p.nx, p.ny, p.gamma, p.boxlen_x, p.boxlen_y, h->stream));
  FVWB_HIP(h, hipMemcpyAsync(dudt, h->stage_w, bytes, hipMemcpyDeviceToHost, h->stream));
  FVWB_HIP(h, hipStreamSynchronize(h->stream));
  return FVWB_OK;
}

extern "C" int fvwb_compute_max_speed(fvwb_handle h, const double* u, double* cmax) {
  if (!h || !u || !cmax) return fail(h, FVWB_ERR_ARG, "NULL argument");
  const fvwb_params& p = h->p;
  const size_t n = (size_t)p.nx * p.ny;
  FVWB_HIP(h, hipMemcpyAsync(h->stage_u, u, 4 * n * sizeof(double), hipMemcpyHostToDevice, h->stream));
  FVWB_HIP(h, hipMemsetAsync(h->red, 0, sizeof(unsigned long long), h->stream));
  FVWB_HIP(h, launch_max_speed_aos(h->stage_u, (long long)n, p.gamma, h->red, h->stream));
  unsigned long long bits = 0;
  FVWB_HIP(h, hipMemcpyAsync(&bits, h->red, sizeof bits, hipMemcpyDeviceToHost, h->stream));
  FVWB_HIP(h, hipStreamSynchronize(h->stream));
  std::memcpy(cmax, &bits, sizeof(double));
  return FVWB_OK;
}

// Reset the step control and reduce the initial wave speed.
static int reset_ctrl(fvwb_ctx* h) {
  const fvwb_params& p = h->p;
  FVWB_HIP(h, hipMemsetAsync(h->ctrl, 0, sizeof(Ctrl), h->stream));
  const double por = p.p0 / p.rho0, igm1 = 1.0 / (p.gamma - 1.0);
  if (p.precision == FVWB_FP64)
    FVWB_HIP(h, launch_cfl_init<double>((const double*)h->buf[0], h->rxc, h->eyc, h->ctrl, p.nx, p.ny, h->pitch,
                                        h->plane, p.gamma, por, igm1, h->stream));
  else
    FVWB_HIP(h, launch_cfl_init<float>((const float*)h->buf[0], h->rxc, h->eyc, h->ctrl, p.nx, p.ny, h->pitch,
                                       h->plane, p.gamma, por, igm1, h->stream));
  h->parity = 0;
  h->uploaded = true;
  return FVWB_OK;
}

extern "C" int fvwb_upload(fvwb_handle h, const double* u, const double* w_eq) {
  if (!h || !u || !w_eq) return fail(h, FVWB_ERR_ARG, "NULL argument");
  const fvwb_params& p = h->p;
  const size_t bytes = 4 * (size_t)p.nx * p.ny * sizeof(double);
  FVWB_HIP(h, hipMemcpyAsync(h->stage_u, u, bytes, hipMemcpyHostToDevice, h->stream));
  FVWB_HIP(h, hipMemcpyAsync(h->w_eq, w_eq, bytes, hipMemcpyHostToDevice, h->stream));
  if (p.precision == FVWB_FP64)
    FVWB_HIP(h, launch_aos_to_dev<double>(h->stage_u, h->w_eq, (double*)h->buf[0], p.nx, p.ny, h->pitch, h->plane,
                                          p.gamma, h->stream));
  else
    FVWB_HIP(h, launch_aos_to_dev<float>(h->stage_u, h->w_eq, (float*)h->buf[0], p.nx, p.ny, h->pitch, h->plane,
                                         p.gamma, h->stream));
  int rc = reset_ctrl(h);
  if (rc) return rc;
  FVWB_HIP(h, hipStreamSynchronize(h->stream));
  return FVWB_OK;
}

extern "C" int fvwb_init_gaussian(fvwb_handle h, double amp, double x0, double y0, double width) {
  if (!h) return fail(h, FVWB_ERR_ARG, "NULL handle");
  const fvwb_params& p = h->p;
  const double dx = p.boxlen_x / p.nx, dy = p.boxlen_y / p.ny, k = p.rho0 * p.g / p.p0;
  if (p.precision == FVWB_FP64)
    FVWB_HIP(h, launch_init_gaussian<double>((double*)h->buf[0], h->w_eq, p.nx, p.ny, h->pitch, h->plane, dx, dy,
                                             p.rho0, p.p0, k, p.gamma, amp, x0, y0, width, h->stream));
  else
    FVWB_HIP(h, launch_init_gaussian<float>((float*)h->buf[0], h->w_eq, p.nx, p.ny, h->pitch, h->plane, dx, dy,
                                            p.rho0, p.p0, k, p.gamma, amp, x0, y0, width, h->stream));
  int rc = reset_ctrl(h);
  if (rc) return rc;
  FVWB_HIP(h, hipStreamSynchronize(h->stream));
  return FVWB_OK;
}

template <typename T>
static int enqueue_steps(fvwb_ctx* h, int nsteps, double t_end) {
  const fvwb_params& p = h->p;
  StageArgs<T> a = base_args<T>(h);
  a.t_end = t_end;
  T* A = (T*)h->buf[0];
  T* B = (T*)h->buf[1];
  T* C = (T*)h->buf[2];
  auto launch = [&](const StageArgs<T>& s, int mode) {
    if constexpr (sizeof(T) == 8)
      return launch_fv2d_stage_f64(s, p.recon, p.flux, p.bc, mode, h->stream);
    else
      return launch_fv2d_stage_f32(s, p.recon, p.flux, p.bc, mode, h->tune.cpl, h->stream);
  };
  for (int n = 0; n < nsteps; ++n) {
    a.parity = h->parity;
    // stage 1: B = A + dt L(A)
    a.in = A; a.un = nullptr; a.out = B; a.first = 1; a.last = 0; a.alpha = 0.0; a.beta = 1.0;
    FVWB_HIP(h, launch(a, 1));
    if (p.integrator == FVWB_RK2_HEUN) {
      // A = 0.5 A + 0.5 (B + dt L(B))   (benchmark_2d.f90:250)
      a.in = B; a.un = A; a.out = A; a.first = 0; a.last = 1; a.alpha = 0.5; a.beta = 0.5;
      FVWB_HIP(h, launch(a, 2));
    } else {
      // C = 3/4 A + 1/4 (B + dt L(B)); A = 1/3 A + 2/3 (C + dt L(C))
      a.in = B; a.un = A; a.out = C; a.first = 0; a.last = 0; a.alpha = 0.75; a.beta = 0.25;
      FVWB_HIP(h, launch(a, 2));
      a.in = C; a.un = A; a.out = A; a.first = 0; a.last = 1; a.alpha = 1.0 / 3.0; a.beta = 2.0 / 3.0;
      FVWB_HIP(h, launch(a, 2));
    }
    h->parity ^= 1;
  }
  return FVWB_OK;
}

extern "C" int fvwb_advance_async(fvwb_handle h, int32_t nsteps, double t_end) {
  if (!h) return fail(h, FVWB_ERR_ARG, "NULL handle");
  if (!h->uploaded) return fail(h, FVWB_ERR_STATE, "advance before upload/init");
  if (nsteps < 0) return fail(h, FVWB_ERR_ARG, "nsteps < 0");
  return h->p.precision == FVWB_FP64 ? enqueue_steps<double>(h, nsteps, t_end)
                                     : enqueue_steps<float>(h, nsteps, t_end);
}

extern "C" int fvwb_get_time(fvwb_handle h, double* t, double* dt_last, int32_t* steps_done) {
  if (!h) return fail(h, FVWB_ERR_ARG, "NULL handle");
  Ctrl c;
  FVWB_HIP(h, hipMemcpyAsync(&c, h->ctrl, sizeof c, hipMemcpyDeviceToHost, h->stream));
  FVWB_HIP(h, hipStreamSynchronize(h->stream));
  if (t) *t = c.t[h->parity];
  if (dt_last) *dt_last = c.dt_last;
  if (steps_done) *steps_done = c.steps_done;
  if (c.err) return fail(h, FVWB_ERR_NONPHYSICAL, "non-physical state (rho<=0, p<=0 or non-finite)");
  return FVWB_OK;
}

extern "C" int fvwb_advance(fvwb_handle h, int32_t nsteps, double t_end, double* t_out, int32_t* steps_done) {
  if (!h) return fail(h, FVWB_ERR_ARG, "NULL handle");
  if (std::isinf(t_end) && t_end > 0) {
    int rc = fvwb_advance_async(h, nsteps, t_end);
    if (rc) return rc;
    return fvwb_get_time(h, t_out, nullptr, steps_done);
  }
  // time-bounded run (while t < t_end, benchmark_2d.f90:238): enqueue chunks and
  // stop once the device time has reached t_end; steps past it are no-ops.
  int32_t left = nsteps;
  double t = 0.0;
  int32_t done = 0;
  while (left > 0) {
    int32_t chunk = left < 64 ? left : 64;
    int rc = fvwb_advance_async(h, chunk, t_end);
    if (rc) return rc;
    left -= chunk;
    rc = fvwb_get_time(h, &t, nullptr, &done);
    if (rc) return rc;
    if (!(t < t_end)) break;
  }
  if (t_out) *t_out = t;
  if (steps_done) *steps_done = done;
  return FVWB_OK;
}

extern "C" int fvwb_download(fvwb_handle h, double* u) {
  if (!h || !u) return fail(h, FVWB_ERR_ARG, "NULL argument");
  if (!h->uploaded) return fail(h, FVWB_ERR_STATE, "download before upload/init");
  const fvwb_params& p = h->p;
  if (p.precision == FVWB_FP64)
    FVWB_HIP(h, launch_dev_to_aos<double>((const double*)h->buf[0], h->w_eq, h->stage_u, p.nx, p.ny, h->pitch,
                                          h->plane, p.gamma, true, h->stream));
  else
    FVWB_HIP(h, launch_dev_to_aos<float>((const float*)h->buf[0], h->w_eq, h->stage_u, p.nx, p.ny, h->pitch,
                                         h->plane, p.gamma, true, h->stream));
  FVWB_HIP(h, hipMemcpyAsync(u, h->stage_u, 4 * (size_t)p.nx * p.ny * sizeof(double), hipMemcpyDeviceToHost,
                             h->stream));
  FVWB_HIP(h, hipStreamSynchronize(h->stream));
  return FVWB_OK;
}

extern "C" int fvwb_download_deviation(fvwb_handle h, double* du) {
  if (!h || !du) return fail(h, FVWB_ERR_ARG, "NULL argument");
  if (!h->uploaded) return fail(h, FVWB_ERR_STATE, "download before upload/init");
  const fvwb_params& p = h->p;
  if (p.precision == FVWB_FP64)
    FVWB_HIP(h, launch_dev_to_aos<double>((const double*)h->buf[0], nullptr, h->stage_u, p.nx, p.ny, h->pitch,
                                          h->plane, p.gamma, false, h->stream));
  else
    FVWB_HIP(h, launch_dev_to_aos<float>((const float*)h->buf[0], nullptr, h->stage_u, p.nx, p.ny, h->pitch,
                                         h->plane, p.gamma, false, h->stream));
  FVWB_HIP(h, hipMemcpyAsync(du, h->stage_u, 4 * (size_t)p.nx * p.ny * sizeof(double), hipMemcpyDeviceToHost,
                             h->stream));
  FVWB_HIP(h, hipStreamSynchronize(h->stream));
  return FVWB_OK;
}

// ---- checkpoint / resume: the deviation planes at t^n, the equilibrium and the clock
extern "C" int fvwb_checkpoint_save(fvwb_handle h, const char* path) {
  if (!h || !path) return fail(h, FVWB_ERR_ARG, "NULL argument");
  if (!h->uploaded) return fail(h, FVWB_ERR_STATE, "checkpoint before upload/init");
  const fvwb_params& p = h->p;
  const fvwb_ckpt::Segment seg[3] = {{h->buf[0], h->elems * h->esize},
                                     {h->w_eq, 4 * (size_t)p.nx * p.ny * sizeof(double)},
                                     {h->ctrl, sizeof(Ctrl)}};
  std::string err;
  const int rc = fvwb_ckpt::save(path, fvwb_ckpt::FV2D, &h->p, sizeof h->p, h->parity, seg, 3, h->stream, err);
  if (rc) h->err = err;
  return rc;
}

extern "C" int fvwb_checkpoint_load(fvwb_handle h, const char* path) {
  if (!h || !path) return fail(h, FVWB_ERR_ARG, "NULL argument");
  const fvwb_params& p = h->p;
  const fvwb_ckpt::Segment seg[3] = {{h->buf[0], h->elems * h->esize},
                                     {h->w_eq, 4 * (size_t)p.nx * p.ny * sizeof(double)},
                                     {h->ctrl, sizeof(Ctrl)}};
  std::string err;
  int32_t parity = 0;
  const int rc = fvwb_ckpt::load(path, fvwb_ckpt::FV2D, &h->p, sizeof h->p, &parity, seg, 3, h->stream, err);
  if (rc) {
    h->uploaded = false;
    h->err = err;
    return rc;
  }
  h->parity = parity;
  h->uploaded = true;
  return FVWB_OK;
}

extern "C" int fvwb_get_stream(fvwb_handle h, void** stream) {
  if (!h || !stream) return fail(h, FVWB_ERR_ARG, "NULL argument");
  *stream = (void*)h->stream;
  return FVWB_OK;
}

namespace fvwb {
hipError_t launch_copy_probe(const void* src, void* dst, long long bytes, hipStream_t s);
}

extern "C" int fvwb_copy_bandwidth(uint64_t bytes, int32_t reps, double* gbps) {
  if (!gbps || bytes < 16 || reps < 1) return fail(nullptr, FVWB_ERR_ARG, "bad copy probe arguments");
  bytes &= ~(uint64_t)15;
  void *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = FVWB_OK;
  float ms = 0.f;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess ||
      hipMemset(a, 1, bytes) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    rc = fail(nullptr, FVWB_ERR_OOM, "copy probe allocation failed");
  } else {
    (void)launch_copy_probe(a, b, (long long)bytes, nullptr);
    (void)hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps; ++r) (void)launch_copy_probe(a, b, (long long)bytes, nullptr);
    (void)hipEventRecord(e1, nullptr);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      rc = fail(nullptr, FVWB_ERR_HIP, "copy probe timing failed");
    else
      *gbps = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

extern "C" int fvwb_set_launch(fvwb_handle h, int32_t rows_per_strip, int32_t columns_per_lane) {
  if (!h) return fail(h, FVWB_ERR_ARG, "NULL handle");
  if (rows_per_strip < 4 || rows_per_strip > 1024 || rows_per_strip % 4 != 0)
    return fail(h, FVWB_ERR_ARG, "rows_per_strip must be a multiple of 4 in [4, 1024]");
  if (columns_per_lane != 2 && columns_per_lane != 4)
    return fail(h, FVWB_ERR_ARG, "columns_per_lane must be 2 or 4");
  h->tune.rows = rows_per_strip;
  if (h->p.precision == FVWB_FP32) h->tune.cpl = columns_per_lane;
  return FVWB_OK;
}

extern "C" int fvwb_get_launch(fvwb_handle h, int32_t* rows_per_strip, int32_t* columns_per_lane) {
  if (!h || !rows_per_strip || !columns_per_lane) return fail(h, FVWB_ERR_ARG, "NULL argument");
  *rows_per_strip = h->tune.rows;
  *columns_per_lane = h->p.precision == FVWB_FP32 ? h->tune.cpl : 2;
  return FVWB_OK;
}
