// fvwb_batch.hip — C ABI of the batched 2D solver (many independent instances).
//
// Each instance is one 2D gravity problem of its own size, gravity strength and
// time step (its own CFL reduction and clock).  All instances share the scheme,
// precision, gamma and cfl.  One persistent launch per Runge-Kutta stage covers
// every strip of every instance; a whole step pair can be replayed from a HIP
// graph.  The launch shape (rows per strip, columns per lane) comes from an
// autotuner that times candidates on a saved copy of the state and caches the
// winner in a small text file.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/fvwb.h"
#include "fvwb_internal.h"

namespace fvwb {
hipError_t launch_fv2d_batch_f64(const BatchArgs<double>& b, int blocks, int recon, int flux, int bc, int mode,
                                 hipStream_t s);
hipError_t launch_fv2d_batch_f32(const BatchArgs<float>& b, int blocks, int recon, int flux, int bc, int mode,
                                 int cpl, hipStream_t s);
template <typename T>
hipError_t launch_aos_to_dev(const double* u, const double* w, T* out, int nx, int ny, int pitch, long long plane,
                             double gamma, hipStream_t s);
template <typename T>
hipError_t launch_dev_to_aos(const T* in, const double* w, double* u, int nx, int ny, int pitch, long long plane,
                             double gamma, bool add_eq, hipStream_t s);
template <typename T>
hipError_t launch_cfl_init(const T* in, const double* rxc, const double* eyc, Ctrl* ctrl, int nx, int ny, int pitch,
                           long long plane, double gamma, double p_over_rho, double inv_gm1, hipStream_t s);
template <typename T>
hipError_t launch_init_gaussian(T* out, double* w, int nx, int ny, int pitch, long long plane, double dx, double dy,
                                double rho0, double p0, double kk, double gamma, double amp, double x0, double y0,
                                double width, hipStream_t s);
}  // namespace fvwb

using namespace fvwb;

struct fvwb_batch {
  int ninst = 0;
  fvwb_params shared{};                 // scheme, precision, gamma, cfl
  std::vector<fvwb_params> p;
  std::vector<long long> off, woff;     // element offsets: state pools, w_eq pool
  // Layout of an instance's block: the four variables' rows of one grid row lie
  // side by side, P = pitch[b] (nx padded to 64) elements each, so element
  // (v, i, j) sits at v P + j (4 P) + i: a row stride of 4P and a variable
  // stride of P.  A strip's four row loads then fall in one 16P-byte span
  // rather than a quarter of the block apart (measured 1.4 % faster in FP32
  // than separate planes on the 512-instance workload, results bit for bit the
  // same).
  std::vector<int> pitch, txo, tyo;
  long long pool_elems = 0, w_elems = 0;
  size_t esize = 4;
  hipStream_t stream = nullptr;
  void* pool[3] = {nullptr, nullptr, nullptr};
  double* w_eq = nullptr;
  double* stage = nullptr;              // AoS staging for one instance
  long long stage_elems = 0;
  double* tabx = nullptr;
  double* taby = nullptr;
  Ctrl* ctrl = nullptr;
  InstDesc* d_inst = nullptr;
  int* d_tile_start = nullptr;
  int ntiles = 0;
  int rows = 8, cpl = 2, blocks = 0;    // launch shape (blocks 0 = automatic)
  int parity = 0;
  bool ready = false;
  bool use_graph = true;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  double graph_tend[2] = {0.0, 0.0};
  std::string err;
};

static thread_local std::string g_batch_err;

#define BATCH_HIP(h, expr)                                                                        \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess) {                                                                       \
      (h)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                              \
      return e_ == hipErrorOutOfMemory ? FVWB_ERR_OOM : FVWB_ERR_HIP;                             \
    }                                                                                             \
  } while (0)

// strides of instance b's block (elements)
static int row_stride(const fvwb_batch* h, int b) { return 4 * h->pitch[b]; }
static long long var_stride(const fvwb_batch* h, int b) { return h->pitch[b]; }

static int bfail(fvwb_batch* h, int code, const std::string& msg) {
  if (h) h->err = msg; else g_batch_err = msg;
  return code;
}

static void drop_graphs(fvwb_batch* h) {
  for (int k = 0; k < 2; ++k)
    if (h->graph[k]) { (void)hipGraphExecDestroy(h->graph[k]); h->graph[k] = nullptr; }
}

// Strip layout for the current launch shape: per-instance descriptors and the
// prefix sums of tiles, uploaded to the device.
static int build_layout(fvwb_batch* h) {
  std::vector<InstDesc> d(h->ninst);
  std::vector<int> start(h->ninst);
  long long total = 0;
  const int cpl = h->shared.precision == FVWB_FP64 ? 2 : h->cpl;
  for (int b = 0; b < h->ninst; ++b) {
    const fvwb_params& p = h->p[b];
    InstDesc& q = d[b];
    q.off = h->off[b];
    q.pitch = row_stride(h, b);
    q.plane = var_stride(h, b);
    q.nx = p.nx; q.ny = p.ny;
    q.rows = h->rows;
    q.ntx = (p.nx + 62 * cpl - 1) / (62 * cpl);
    q.ntiles = q.ntx * ((p.ny + h->rows - 1) / h->rows);
    q.tx_off = h->txo[b]; q.ty_off = h->tyo[b];
    const double dx = p.boxlen_x / p.nx, dy = p.boxlen_y / p.ny;
    q.h = dx < dy ? dx : dy;
    q.cfl = p.cfl; q.gamma = p.gamma; q.g = p.g;
    q.p_over_rho = p.p0 / p.rho0; q.inv_gm1 = 1.0 / (p.gamma - 1.0);
    q.odx = 1.0 / dx; q.ody = 1.0 / dy;
    start[b] = (int)total;
    total += q.ntiles;
  }
  if (total > (1ll << 30)) return bfail(h, FVWB_ERR_ARG, "too many strips in the batch");
  h->ntiles = (int)total;
  BATCH_HIP(h, hipMemcpyAsync(h->d_inst, d.data(), sizeof(InstDesc) * h->ninst, hipMemcpyHostToDevice, h->stream));
  BATCH_HIP(h, hipMemcpyAsync(h->d_tile_start, start.data(), sizeof(int) * h->ninst, hipMemcpyHostToDevice, h->stream));
  BATCH_HIP(h, hipStreamSynchronize(h->stream));
  drop_graphs(h);
  return FVWB_OK;
}

extern "C" int fvwb_batch_create(fvwb_batch_handle* out, int32_t ninst, const fvwb_params* per_inst) {
  if (!out || !per_inst || ninst < 1) return bfail(nullptr, FVWB_ERR_ARG, "bad arguments");
  *out = nullptr;
  const fvwb_params& s0 = per_inst[0];
  for (int b = 0; b < ninst; ++b) {
    const fvwb_params& p = per_inst[b];
    if (p.nvar != 4 || p.nx < 4 || p.ny < 4 || (long long)p.nx * p.ny > (1ll << 28))
      return bfail(nullptr, FVWB_ERR_ARG, "instance " + std::to_string(b) + ": bad size (4 <= nx,ny; nx*ny <= 2^28)");
    if (p.bc != s0.bc || p.recon != s0.recon || p.flux != s0.flux || p.integrator != s0.integrator ||
        p.precision != s0.precision || p.gamma != s0.gamma || p.cfl != s0.cfl)
      return bfail(nullptr, FVWB_ERR_ARG, "instance " + std::to_string(b) + ": scheme differs from instance 0");
    if (!(p.gamma > 1.0) || !(p.boxlen_x > 0) || !(p.boxlen_y > 0) || !(p.cfl > 0) || !(p.rho0 > 0) ||
        !(p.p0 > 0) || !(p.g >= 0))
      return bfail(nullptr, FVWB_ERR_ARG, "instance " + std::to_string(b) + ": non-physical constants");
  }
  if ((s0.bc != 0 && s0.bc != 1) || (s0.recon != 0 && s0.recon != 1) || (s0.flux != 0 && s0.flux != 1) ||
      (s0.integrator != 0 && s0.integrator != 1) || (s0.precision != 0 && s0.precision != 1))
    return bfail(nullptr, FVWB_ERR_ARG, "unknown scheme selector");
  fvwb_batch* h = new (std::nothrow) fvwb_batch();
  if (!h) return bfail(nullptr, FVWB_ERR_OOM, "host allocation failed");
  h->ninst = ninst;
  h->shared = s0;
  h->p.assign(per_inst, per_inst + ninst);
  h->esize = s0.precision == FVWB_FP64 ? 8 : 4;
  long long tx = 0, ty = 0;
  for (int b = 0; b < ninst; ++b) {
    const fvwb_params& p = per_inst[b];
    const int pitch = (p.nx + 63) / 64 * 64;
    h->pitch.push_back(pitch);
    h->off.push_back(h->pool_elems);
    h->pool_elems += 4ll * pitch * p.ny;
    h->woff.push_back(h->w_elems);
    h->w_elems += 4ll * p.nx * p.ny;
    h->stage_elems = std::max(h->stage_elems, 4ll * p.nx * p.ny);
    h->txo.push_back((int)tx); tx += 2ll * p.nx + 1;
    h->tyo.push_back((int)ty); ty += 2ll * p.ny + 1;
  }
  if (tx > (1ll << 31) || ty > (1ll << 31)) { delete h; return bfail(nullptr, FVWB_ERR_ARG, "batch too large"); }
  hipError_t e = hipStreamCreate(&h->stream);
  if (e != hipSuccess) { delete h; return bfail(nullptr, FVWB_ERR_HIP, hipGetErrorString(e)); }
  auto alloc = [&](void** ptr, size_t bytes) {
    return hipMalloc(ptr, bytes) == hipSuccess;
  };
  bool ok = true;
  for (int k = 0; k < 3; ++k) ok = ok && alloc(&h->pool[k], (size_t)h->pool_elems * h->esize);
  ok = ok && alloc((void**)&h->w_eq, (size_t)h->w_elems * sizeof(double));
  ok = ok && alloc((void**)&h->stage, (size_t)h->stage_elems * sizeof(double));
  ok = ok && alloc((void**)&h->tabx, (size_t)tx * sizeof(double));
  ok = ok && alloc((void**)&h->taby, (size_t)ty * sizeof(double));
  ok = ok && alloc((void**)&h->ctrl, sizeof(Ctrl) * ninst);
  ok = ok && alloc((void**)&h->d_inst, sizeof(InstDesc) * ninst);
  ok = ok && alloc((void**)&h->d_tile_start, sizeof(int) * ninst);
  if (!ok) {
    fvwb_batch_destroy(h);
    return bfail(nullptr, FVWB_ERR_OOM, "device allocation failed");
  }
  // separable FP64 equilibrium tables per instance (as in the single solver)
  std::vector<double> hx(tx), hy(ty);
  for (int b = 0; b < ninst; ++b) {
    const fvwb_params& p = per_inst[b];
    const double dx = p.boxlen_x / p.nx, dy = p.boxlen_y / p.ny, k = p.rho0 * p.g / p.p0;
    double* rxf = hx.data() + h->txo[b];
    double* rxc = rxf + p.nx + 1;
    for (int i = 0; i <= p.nx; ++i) rxf[i] = p.rho0 * std::exp(-(k * ((double)i * dx)));
    for (int i = 0; i < p.nx; ++i) rxc[i] = p.rho0 * std::exp(-(k * (((double)i + 0.5) * dx)));
    double* eyc = hy.data() + h->tyo[b];
    double* eyf = eyc + p.ny;
    for (int j = 0; j < p.ny; ++j) eyc[j] = std::exp(-(k * (((double)j + 0.5) * dy)));
    for (int j = 0; j <= p.ny; ++j) eyf[j] = std::exp(-(k * ((double)j * dy)));
  }
  if (hipMemcpy(h->tabx, hx.data(), tx * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->taby, hy.data(), ty * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(h->ctrl, 0, sizeof(Ctrl) * ninst) != hipSuccess) {
    fvwb_batch_destroy(h);
    return bfail(nullptr, FVWB_ERR_HIP, "table upload failed");
  }
  // default launch shape: short strips (many waves), FP32 lanes own 2 columns
  h->rows = 8;
  h->cpl = 2;
  int rc = build_layout(h);
  if (rc) { g_batch_err = h->err; fvwb_batch_destroy(h); return rc; }
  *out = h;
  return FVWB_OK;
}

extern "C" int fvwb_batch_destroy(fvwb_batch_handle h) {
  if (!h) return FVWB_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  drop_graphs(h);
  for (int k = 0; k < 3; ++k) if (h->pool[k]) (void)hipFree(h->pool[k]);
  void* ps[] = {h->w_eq, h->stage, h->tabx, h->taby, h->ctrl, h->d_inst, h->d_tile_start};
  for (void* q : ps) if (q) (void)hipFree(q);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return FVWB_OK;
}

extern "C" const char* fvwb_batch_last_error(fvwb_batch_handle h) {
  return h ? h->err.c_str() : g_batch_err.c_str();
}

static int check_inst(fvwb_batch* h, int inst) {
  if (!h) return bfail(h, FVWB_ERR_ARG, "NULL handle");
  if (inst < 0 || inst >= h->ninst) return bfail(h, FVWB_ERR_ARG, "instance index out of range");
  return FVWB_OK;
}

// Initial reduction of max(|v|+c) of instance b into its cmax[0] slots.
static int cfl_init(fvwb_batch* h, int b) {
  const fvwb_params& p = h->p[b];
  const double por = p.p0 / p.rho0, igm1 = 1.0 / (p.gamma - 1.0);
  const double* rxc = h->tabx + h->txo[b] + p.nx + 1;
  const double* eyc = h->taby + h->tyo[b];
  if (h->esize == 8)
    BATCH_HIP(h, launch_cfl_init<double>((const double*)h->pool[0] + h->off[b], rxc, eyc, h->ctrl + b, p.nx, p.ny,
                                         row_stride(h, b), var_stride(h, b), p.gamma, por, igm1, h->stream));
  else
    BATCH_HIP(h, launch_cfl_init<float>((const float*)h->pool[0] + h->off[b], rxc, eyc, h->ctrl + b, p.nx, p.ny,
                                        row_stride(h, b), var_stride(h, b), p.gamma, por, igm1, h->stream));
  return FVWB_OK;
}

// All instances restart at t = 0 with the current state (pool A).
static int restart_all(fvwb_batch* h) {
  BATCH_HIP(h, hipMemsetAsync(h->ctrl, 0, sizeof(Ctrl) * h->ninst, h->stream));
  for (int b = 0; b < h->ninst; ++b) {
    int rc = cfl_init(h, b);
    if (rc) return rc;
  }
  h->parity = 0;
  h->ready = true;
  drop_graphs(h);
  BATCH_HIP(h, hipStreamSynchronize(h->stream));
  return FVWB_OK;
}

extern "C" int fvwb_batch_init_gaussian(fvwb_batch_handle h, const double* amp, double x0, double y0, double width) {
  if (!h || !amp) return bfail(h, FVWB_ERR_ARG, "NULL argument");
  for (int b = 0; b < h->ninst; ++b) {
    const fvwb_params& p = h->p[b];
    const double dx = p.boxlen_x / p.nx, dy = p.boxlen_y / p.ny, k = p.rho0 * p.g / p.p0;
      if (h->esize == 8)
      BATCH_HIP(h, launch_init_gaussian<double>((double*)h->pool[0] + h->off[b], h->w_eq + h->woff[b], p.nx, p.ny,
                                                row_stride(h, b), var_stride(h, b), dx, dy, p.rho0, p.p0, k, p.gamma,
                                                amp[b], x0, y0, width, h->stream));
    else
      BATCH_HIP(h, launch_init_gaussian<float>((float*)h->pool[0] + h->off[b], h->w_eq + h->woff[b], p.nx, p.ny,
                                               row_stride(h, b), var_stride(h, b), dx, dy, p.rho0, p.p0, k, p.gamma,
                                               amp[b], x0, y0, width, h->stream));
  }
  return restart_all(h);
}

extern "C" int fvwb_batch_upload(fvwb_batch_handle h, int32_t inst, const double* u, const double* w_eq) {
  int rc = check_inst(h, inst);
  if (rc) return rc;
  if (!u || !w_eq) return bfail(h, FVWB_ERR_ARG, "NULL argument");
  const fvwb_params& p = h->p[inst];
  const size_t bytes = 4 * (size_t)p.nx * p.ny * sizeof(double);
  BATCH_HIP(h, hipMemcpyAsync(h->stage, u, bytes, hipMemcpyHostToDevice, h->stream));
  BATCH_HIP(h, hipMemcpyAsync(h->w_eq + h->woff[inst], w_eq, bytes, hipMemcpyHostToDevice, h->stream));
  if (h->esize == 8)
    BATCH_HIP(h, launch_aos_to_dev<double>(h->stage, h->w_eq + h->woff[inst], (double*)h->pool[0] + h->off[inst],
                                           p.nx, p.ny, row_stride(h, inst), var_stride(h, inst), p.gamma, h->stream));
  else
    BATCH_HIP(h, launch_aos_to_dev<float>(h->stage, h->w_eq + h->woff[inst], (float*)h->pool[0] + h->off[inst],
                                          p.nx, p.ny, row_stride(h, inst), var_stride(h, inst), p.gamma, h->stream));
  BATCH_HIP(h, hipStreamSynchronize(h->stream));
  h->ready = false;  // clocks restart at the next restart (fvwb_batch_reset_time)
  return FVWB_OK;
}

extern "C" int fvwb_batch_reset_time(fvwb_batch_handle h) {
  if (!h) return bfail(h, FVWB_ERR_ARG, "NULL handle");
  return restart_all(h);
}

template <typename T>
static int enqueue_batch_steps(fvwb_batch* h, int nsteps, double t_end) {
  const fvwb_params& s = h->shared;
  BatchArgs<T> a{};
  a.inst = h->d_inst; a.tile_start = h->d_tile_start; a.ninst = h->ninst; a.ntiles = h->ntiles;
  a.tabx = h->tabx; a.taby = h->taby; a.ctrl = h->ctrl; a.t_end = t_end;
  T* A = (T*)h->pool[0];
  T* B = (T*)h->pool[1];
  T* C = (T*)h->pool[2];
  int blocks = h->blocks > 0 ? h->blocks : (h->ntiles + 3) / 4;
  blocks = std::max(1, std::min(blocks, (h->ntiles + 3) / 4));
  auto launch = [&](const BatchArgs<T>& b, int mode) {
    if constexpr (sizeof(T) == 8)
      return launch_fv2d_batch_f64(b, blocks, s.recon, s.flux, s.bc, mode, h->stream);
    else
      return launch_fv2d_batch_f32(b, blocks, s.recon, s.flux, s.bc, mode, h->cpl, h->stream);
  };
  for (int n = 0; n < nsteps; ++n) {
    a.parity = h->parity;
    a.in = A; a.un = nullptr; a.out = B; a.first = 1; a.last = 0; a.alpha = 0.0; a.beta = 1.0;
    BATCH_HIP(h, launch(a, 1));
    if (s.integrator == FVWB_RK2_HEUN) {
      a.in = B; a.un = A; a.out = A; a.first = 0; a.last = 1; a.alpha = 0.5; a.beta = 0.5;
      BATCH_HIP(h, launch(a, 2));
    } else {
      a.in = B; a.un = A; a.out = C; a.first = 0; a.last = 0; a.alpha = 0.75; a.beta = 0.25;
      BATCH_HIP(h, launch(a, 2));
      a.in = C; a.un = A; a.out = A; a.first = 0; a.last = 1; a.alpha = 1.0 / 3.0; a.beta = 2.0 / 3.0;
      BATCH_HIP(h, launch(a, 2));
    }
    h->parity ^= 1;
  }
  return FVWB_OK;
}

static int enqueue_steps(fvwb_batch* h, int nsteps, double t_end) {
  return h->esize == 8 ? enqueue_batch_steps<double>(h, nsteps, t_end) : enqueue_batch_steps<float>(h, nsteps, t_end);
}

// A graph holding two steps (both slot parities) starting at parity `p`.
static int graph_for(fvwb_batch* h, int p, double t_end, hipGraphExec_t* out) {
  if (h->graph[p] && h->graph_tend[p] == t_end) { *out = h->graph[p]; return FVWB_OK; }
  if (h->graph[p]) { (void)hipGraphExecDestroy(h->graph[p]); h->graph[p] = nullptr; }
  const int saved = h->parity;
  h->parity = p;
  hipGraph_t g = nullptr;
  BATCH_HIP(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_steps(h, 2, t_end);
  hipError_t e = hipStreamEndCapture(h->stream, &g);
  h->parity = saved;
  if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
  if (e != hipSuccess) return bfail(h, FVWB_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(&h->graph[p], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) { h->graph[p] = nullptr; return bfail(h, FVWB_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e)); }
  h->graph_tend[p] = t_end;
  *out = h->graph[p];
  return FVWB_OK;
}

extern "C" int fvwb_batch_advance_async(fvwb_batch_handle h, int32_t nsteps, double t_end) {
  if (!h) return bfail(h, FVWB_ERR_ARG, "NULL handle");
  if (!h->ready) return bfail(h, FVWB_ERR_STATE, "advance before init / reset_time");
  if (nsteps < 0) return bfail(h, FVWB_ERR_ARG, "nsteps < 0");
  if (!h->use_graph) return enqueue_steps(h, nsteps, t_end);
  int pairs = nsteps / 2;
  if (pairs > 0) {
    hipGraphExec_t g = nullptr;
    int rc = graph_for(h, h->parity, t_end, &g);
    if (rc) return rc;
    for (int k = 0; k < pairs; ++k) BATCH_HIP(h, hipGraphLaunch(g, h->stream));
  }
  if (nsteps & 1) return enqueue_steps(h, 1, t_end);
  return FVWB_OK;
}

extern "C" int fvwb_batch_advance(fvwb_batch_handle h, int32_t nsteps, double t_end) {
  int rc = fvwb_batch_advance_async(h, nsteps, t_end);
  if (rc) return rc;
  BATCH_HIP(h, hipStreamSynchronize(h->stream));
  std::vector<Ctrl> c(h->ninst);
  BATCH_HIP(h, hipMemcpy(c.data(), h->ctrl, sizeof(Ctrl) * h->ninst, hipMemcpyDeviceToHost));
  for (int b = 0; b < h->ninst; ++b)
    if (c[b].err) return bfail(h, FVWB_ERR_NONPHYSICAL, "instance " + std::to_string(b) + ": non-physical state");
  return FVWB_OK;
}

extern "C" int fvwb_batch_get_time(fvwb_batch_handle h, int32_t inst, double* t, double* dt_last, int32_t* steps) {
  int rc = check_inst(h, inst);
  if (rc) return rc;
  Ctrl c;
  BATCH_HIP(h, hipMemcpyAsync(&c, h->ctrl + inst, sizeof c, hipMemcpyDeviceToHost, h->stream));
  BATCH_HIP(h, hipStreamSynchronize(h->stream));
  if (t) *t = c.t[h->parity];
  if (dt_last) *dt_last = c.dt_last;
  if (steps) *steps = c.steps_done;
  return c.err ? bfail(h, FVWB_ERR_NONPHYSICAL, "non-physical state") : FVWB_OK;
}

static int download_impl(fvwb_batch* h, int inst, double* u, bool add_eq) {
  int rc = check_inst(h, inst);
  if (rc) return rc;
  if (!u) return bfail(h, FVWB_ERR_ARG, "NULL argument");
  const fvwb_params& p = h->p[inst];
  if (h->esize == 8)
    BATCH_HIP(h, launch_dev_to_aos<double>((const double*)h->pool[0] + h->off[inst], h->w_eq + h->woff[inst],
                                           h->stage, p.nx, p.ny, row_stride(h, inst), var_stride(h, inst), p.gamma, add_eq,
                                           h->stream));
  else
    BATCH_HIP(h, launch_dev_to_aos<float>((const float*)h->pool[0] + h->off[inst], h->w_eq + h->woff[inst],
                                          h->stage, p.nx, p.ny, row_stride(h, inst), var_stride(h, inst), p.gamma, add_eq,
                                          h->stream));
  BATCH_HIP(h, hipMemcpyAsync(u, h->stage, 4 * (size_t)p.nx * p.ny * sizeof(double), hipMemcpyDeviceToHost,
                              h->stream));
  BATCH_HIP(h, hipStreamSynchronize(h->stream));
  return FVWB_OK;
}

extern "C" int fvwb_batch_download(fvwb_batch_handle h, int32_t inst, double* u) {
  return download_impl(h, inst, u, true);
}

extern "C" int fvwb_batch_download_deviation(fvwb_batch_handle h, int32_t inst, double* du) {
  return download_impl(h, inst, du, false);
}

extern "C" int fvwb_batch_get_stream(fvwb_batch_handle h, void** stream) {
  if (!h || !stream) return bfail(h, FVWB_ERR_ARG, "NULL argument");
  *stream = (void*)h->stream;
  return FVWB_OK;
}

extern "C" int fvwb_batch_use_graph(fvwb_batch_handle h, int32_t enable) {
  if (!h) return bfail(h, FVWB_ERR_ARG, "NULL handle");
  h->use_graph = enable != 0;
  drop_graphs(h);
  return FVWB_OK;
}

extern "C" int fvwb_batch_set_launch(fvwb_batch_handle h, int32_t rows, int32_t cpl, int32_t blocks) {
  if (!h) return bfail(h, FVWB_ERR_ARG, "NULL handle");
  if (rows < 4 || rows > 1024 || rows % 4 != 0) return bfail(h, FVWB_ERR_ARG, "rows must be a multiple of 4 in [4,1024]");
  if (cpl != 2 && cpl != 4) return bfail(h, FVWB_ERR_ARG, "cpl must be 2 or 4");
  if (blocks < 0) return bfail(h, FVWB_ERR_ARG, "blocks < 0");
  h->rows = rows;
  if (h->shared.precision == FVWB_FP32) h->cpl = cpl;
  h->blocks = blocks;
  return build_layout(h);
}

extern "C" int fvwb_batch_get_launch(fvwb_batch_handle h, int32_t* rows, int32_t* cpl, int32_t* blocks,
                                     int32_t* ntiles) {
  if (!h) return bfail(h, FVWB_ERR_ARG, "NULL handle");
  if (rows) *rows = h->rows;
  if (cpl) *cpl = h->shared.precision == FVWB_FP32 ? h->cpl : 2;
  if (blocks) *blocks = h->blocks;
  if (ntiles) *ntiles = h->ntiles;
  return FVWB_OK;
}

// ---- autotuner ------------------------------------------------------------------

static unsigned long long batch_key(const fvwb_batch* h) {
  unsigned long long x = 1469598103934665603ull;  // FNV-1a over the shape of the batch
  auto mix = [&](unsigned long long v) {
    for (int k = 0; k < 8; ++k) { x ^= (v >> (8 * k)) & 0xff; x *= 1099511628211ull; }
  };
  mix(h->ninst);
  mix(h->shared.precision); mix(h->shared.recon); mix(h->shared.flux); mix(h->shared.bc); mix(h->shared.integrator);
  for (const fvwb_params& p : h->p) { mix(p.nx); mix(p.ny); }
  return x;
}

static bool cache_lookup(const char* path, unsigned long long key, int* rows, int* cpl, int* blocks) {
  if (!path) return false;
  FILE* f = std::fopen(path, "r");
  if (!f) return false;
  char line[256];
  bool found = false;
  while (std::fgets(line, sizeof line, f)) {
    unsigned long long k;
    int r, c, b;
    if (std::sscanf(line, "fv2d-batch %llx rows=%d cpl=%d blocks=%d", &k, &r, &c, &b) == 4 && k == key) {
      *rows = r; *cpl = c; *blocks = b; found = true;
    }
  }
  std::fclose(f);
  return found;
}

static void cache_store(const char* path, unsigned long long key, int rows, int cpl, int blocks, double ms) {
  if (!path) return;
  FILE* f = std::fopen(path, "a");
  if (!f) return;
  std::fprintf(f, "fv2d-batch %016llx rows=%d cpl=%d blocks=%d ms_per_step=%.4f\n", key, rows, cpl, blocks, ms);
  std::fclose(f);
}

extern "C" int fvwb_batch_autotune(fvwb_batch_handle h, int32_t trial_steps, const char* cache_path,
                                   int32_t* rows_out, int32_t* cpl_out) {
  if (!h) return bfail(h, FVWB_ERR_ARG, "NULL handle");
  if (!h->ready) return bfail(h, FVWB_ERR_STATE, "autotune before init / reset_time");
  if (trial_steps < 1) trial_steps = 2;
  const unsigned long long key = batch_key(h);
  int rows = h->rows, cpl = h->cpl, blocks = h->blocks;
  if (!cache_lookup(cache_path, key, &rows, &cpl, &blocks)) {
    // save the state and clocks, time each candidate, restore
    const size_t bytes = (size_t)h->pool_elems * h->esize;
    void* saved = nullptr;
    std::vector<Ctrl> ctrl(h->ninst);
    BATCH_HIP(h, hipMalloc(&saved, bytes));
    BATCH_HIP(h, hipMemcpyAsync(saved, h->pool[0], bytes, hipMemcpyDeviceToDevice, h->stream));
    BATCH_HIP(h, hipMemcpyAsync(ctrl.data(), h->ctrl, sizeof(Ctrl) * h->ninst, hipMemcpyDeviceToHost, h->stream));
    BATCH_HIP(h, hipStreamSynchronize(h->stream));
    const int saved_parity = h->parity;
    hipEvent_t e0, e1;
    BATCH_HIP(h, hipEventCreate(&e0));
    BATCH_HIP(h, hipEventCreate(&e1));
    const int row_opts[] = {4, 8, 12, 16, 24, 32};
    // FP64 lanes always own 2 columns; FP32 lanes own 2 or 4 (one column per
    // lane measured no faster on MI355X)
    const int cpl_opts32[] = {2, 4}, cpl_opts64[] = {2};
    const int* cpl_opts = h->shared.precision == FVWB_FP32 ? cpl_opts32 : cpl_opts64;
    const int ncpl = h->shared.precision == FVWB_FP32 ? 2 : 1;
    double best = 1e300;
    int rc = FVWB_OK;
    for (int ci = 0; ci < ncpl && rc == FVWB_OK; ++ci)
      for (int r : row_opts) {
        h->rows = r; h->cpl = cpl_opts[ci]; h->blocks = 0;
        rc = build_layout(h);
        if (rc) break;
        rc = enqueue_steps(h, 1, INFINITY);  // warm the kernel
        if (rc) break;
        (void)hipEventRecord(e0, h->stream);
        rc = enqueue_steps(h, trial_steps, INFINITY);
        (void)hipEventRecord(e1, h->stream);
        if (rc) break;
        float ms = 0.f;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
          rc = bfail(h, FVWB_ERR_HIP, "autotune timing failed");
          break;
        }
        if (ms < best) { best = ms; rows = r; cpl = cpl_opts[ci]; blocks = 0; }
        // restore the saved state before the next candidate
        if (hipMemcpyAsync(h->pool[0], saved, bytes, hipMemcpyDeviceToDevice, h->stream) != hipSuccess ||
            hipMemcpyAsync(h->ctrl, ctrl.data(), sizeof(Ctrl) * h->ninst, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
            hipStreamSynchronize(h->stream) != hipSuccess) {
          rc = bfail(h, FVWB_ERR_HIP, "autotune restore failed");
          break;
        }
        h->parity = saved_parity;
      }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(saved);
    if (rc) return rc;
    cache_store(cache_path, key, rows, cpl, blocks, best / trial_steps);
  }
  h->rows = rows; h->cpl = cpl; h->blocks = blocks;
  int rc = build_layout(h);
  if (rc) return rc;
  if (rows_out) *rows_out = rows;
  if (cpl_out) *cpl_out = h->shared.precision == FVWB_FP32 ? cpl : 2;
  return FVWB_OK;
}
