// fvwb_rd.hip — the reference's 2D residual distribution (benchmark_rd.f90,
// solver 'RD1') on MI355X.
//
// compute_update_rd (:612-745) scatters four corner residuals of every quad
// cell into its nodes.  Here each workgroup computes the corner values of the
// cells around a tile of nodes into LDS and every node gathers its own, adding
// them in the order the reference's cell loop (x outer, y inner) adds them, so
// the sums round identically without atomics.  The two boundary passes apply
// the reference's copies (:722-737) in their order.
// FP64 with contraction off and the reference's expression order throughout:
// the results equal the oracle's bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>

#include "../../include/fvwb.h"
#include "fvwb_checkpoint.h"

#pragma clang fp contract(off)

namespace rd {

struct Geo {
  int nx, ny, bc;
  double gamma, dx, dy;
};

__device__ __forceinline__ void primitive(double gamma, const double u[4], double w[4]) {
  w[0] = u[0];
  w[1] = u[1] / u[0];
  w[2] = u[2] / u[0];
  w[3] = (gamma - 1.0) * (u[3] - 0.5 * w[0] * (w[1] * w[1] + w[2] * w[2]));
}

// compute_speed (:478-493) of the vector it is handed
__device__ __forceinline__ double speed(double gamma, const double u[4]) {
  double w[4];
  primitive(gamma, u, w);
  const double cs = sqrt(gamma * fmax(w[3], 1e-10) / fmax(w[0], 1e-10));
  return sqrt(w[1] * w[1] + w[2] * w[2]) + cs;
}

__device__ __forceinline__ void load4(const double* a, int nx, int i, int j, double v[4]) {
  const double4 q = *reinterpret_cast<const double4*>(a + 4 * ((long long)j * nx + i));
  v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
}

// Corner residuals of cell (i, j) (nodes i..i+1, j..j+1): out[k*4 + v] for
// corners k = 11, 12, 21, 22 (benchmark_rd.f90:653-697).
__device__ __forceinline__ void cell_phi(const Geo& g, const double* __restrict__ u,
                                         const double* __restrict__ diff, double dt, int i, int j,
                                         double out[16]) {
  double uu[4][4], ww[4][4], ff[4][4], gg[4][4], ss[4][4], sp[4];
  load4(u, g.nx, i, j, uu[0]);
  load4(u, g.nx, i + 1, j, uu[1]);
  load4(u, g.nx, i, j + 1, uu[2]);
  load4(u, g.nx, i + 1, j + 1, uu[3]);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    // compute_flux (:549-571): ff along x, gg along y
    double w[4];
    primitive(g.gamma, uu[k], w);
    gg[k][0] = w[0] * w[2];
    gg[k][1] = w[0] * w[1] * w[2];
    gg[k][2] = w[2] * uu[k][2] + w[3];
    gg[k][3] = w[2] * uu[k][3] + w[2] * w[3];
    ff[k][0] = w[0] * w[1];
    ff[k][1] = w[1] * uu[k][1] + w[3];
    ff[k][2] = w[0] * w[1] * w[2];
    ff[k][3] = w[1] * uu[k][3] + w[1] * w[3];
    primitive(g.gamma, uu[k], ww[k]);
    // get_source (:747-767), grad_p = (1, 1): it receives the primitive
    // vector (:681-684) and converts it once more (:756), so its velocities
    // are ww[1]/ww[0] and ww[2]/ww[0]
    double w2[4];
    primitive(g.gamma, ww[k], w2);
    ss[k][0] = 0.0;
    ss[k][1] = -w2[0] * 1.0;
    ss[k][2] = -w2[0] * 1.0;
    ss[k][3] = -w2[0] * (w2[1] * 1.0 + w2[2] * 1.0);
    // the reference hands compute_speed the primitive vector (:692-695)
    sp[k] = speed(g.gamma, ww[k]);
  }
  double alpha = sp[0];
#pragma unroll
  for (int k = 1; k < 4; ++k)
    if (sp[k] > alpha) alpha = sp[k];
  double dd[4][4];
  load4(diff, g.nx, i, j, dd[0]);
  load4(diff, g.nx, i + 1, j, dd[1]);
  load4(diff, g.nx, i, j + 1, dd[2]);
  load4(diff, g.nx, i + 1, j + 1, dd[3]);
  const double dx = g.dx, dy = g.dy;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    // :673-678, corners 11, 12, 21, 22 = uu[0..3]
    const double totalphi = -0.5 * (ff[0][v] + ff[2][v]) + 0.5 * (ff[1][v] + ff[3][v]) -
                            0.5 * (gg[3][v] + gg[2][v]) + 0.5 * (gg[0][v] + gg[1][v]) -
                            0.25 * (ss[0][v] + ss[2][v] + ss[1][v] + ss[3][v]) * dx * dy;
    const double ubar = 0.25 * (uu[0][v] + uu[3][v] + uu[2][v] + uu[1][v]);
    // :688-697
    out[0 * 4 + v] = dy * dx * dd[0][v] * 0.25 + dt * (totalphi * 0.25 + alpha * (uu[0][v] - ubar) * dy);
    out[1 * 4 + v] = dy * dx * dd[1][v] * 0.25 + dt * (totalphi * 0.25 + alpha * (uu[1][v] - ubar) * dy);
    out[2 * 4 + v] = dy * dx * dd[2][v] * 0.25 + dt * (totalphi * 0.25 + alpha * (uu[2][v] - ubar) * dx);
    out[3 * 4 + v] = dy * dx * dd[3][v] * 0.25 + dt * (totalphi * 0.25 + alpha * (uu[3][v] - ubar) * dx);
  }
}

// The resident time loop's clock, on the device so that steps follow each
// other without a host round trip: dt of the current step, whether it runs
// (t < t_end and no failure), and the running totals.
struct Clock {
  double t, dt, dt_last;
  long long steps;
  int active, err;
};

// One workgroup per tile of TX x TY nodes: the corner residuals of the
// (TX+1) x (TY+1) cells that touch the tile (its cells and a one-cell halo,
// which the neighbouring tiles compute as well) go to LDS, then every node
// gathers its up to four corner values from there in the order the
// reference's x-outer, y-inner cell loop adds them (cells (a-1,b-1), (a-1,b),
// (a,b-1), (a,b)), so the sums round as the reference's do and no corner value
// travels through HBM.
// 31 x 7 nodes make the 32 x 8 = 256 cells one per thread (one fill round).
constexpr int TX = 31, TY = 7, CX = TX + 1, CY = TY + 1, NC = CX * CY;
static_assert(NC == 256, "one cell per thread of the workgroup");

// dt is the argument, or the clock's when ck is given (the resident loop).
__global__ __launch_bounds__(256) void rd_update_kernel(Geo g, const double* __restrict__ u,
                                                        const double* __restrict__ diff, double dt,
                                                        const Clock* __restrict__ ck, double* __restrict__ dudt) {
  if (ck) {
    if (!ck->active) return;
    dt = ck->dt;
  }
  __shared__ double lds[16 * NC];           // [corner*4 + v][cell], 32 KB
  const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY;
  const int c = threadIdx.x;                // cell (x0-1 + c%CX, y0-1 + c/CX)
  const int ci = x0 - 1 + c % CX, cj = y0 - 1 + c / CX;
  if (ci >= 0 && ci < g.nx - 1 && cj >= 0 && cj < g.ny - 1) {
    double phi[16];
    cell_phi(g, u, diff, dt, ci, cj, phi);
#pragma unroll
    for (int q = 0; q < 16; ++q) lds[q * NC + c] = phi[q];
  }
  __syncthreads();
  const int tx = threadIdx.x % CX, ty = threadIdx.x / CX;   // node (x0+tx, y0+ty)
  const int a = x0 + tx, b = y0 + ty;
  if (tx >= TX || ty >= TY || a >= g.nx || b >= g.ny) return;
  const int cx = g.nx - 1, cy = g.ny - 1;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  // local cell index of cell (a-1+di, b-1+dj) is (ty+dj)*CX + tx+di
  auto add = [&](int di, int dj, int corner) {
    const int c = (ty + dj) * CX + tx + di;
#pragma unroll
    for (int v = 0; v < 4; ++v) s[v] = s[v] + lds[(corner * 4 + v) * NC + c];
  };
  if (a >= 1 && b >= 1) add(0, 0, 3);
  if (a >= 1 && b < cy) add(0, 1, 1);
  if (a < cx && b >= 1) add(1, 0, 2);
  if (a < cx && b < cy) add(1, 1, 0);
  *reinterpret_cast<double4*>(dudt + 4 * ((long long)b * g.nx + a)) = double4{s[0], s[1], s[2], s[3]};
}

__device__ __forceinline__ void copy_node(double* d, int nx, int ti, int tj, int si, int sj) {
  double* t = d + 4 * ((long long)tj * nx + ti);
  const double* s = d + 4 * ((long long)sj * nx + si);
  t[0] = s[0]; t[1] = s[1]; t[2] = s[2]; t[3] = s[3];
}

// :722-737, the copies along x (one thread per row j), then along y (one per column i)
__global__ void rd_bc_x_kernel(Geo g, double* __restrict__ d) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= g.ny) return;
  if (g.bc == FVWB_RD_BC_PERIODIC) {
    copy_node(d, g.nx, 0, j, g.nx - 1, j);
    copy_node(d, g.nx, g.nx - 1, j, 0, j);
  } else {
    copy_node(d, g.nx, 0, j, 1, j);
    copy_node(d, g.nx, g.nx - 1, j, g.nx - 2, j);
  }
}
__global__ void rd_bc_y_kernel(Geo g, double* __restrict__ d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.nx) return;
  if (g.bc == FVWB_RD_BC_PERIODIC) {
    copy_node(d, g.nx, i, g.ny - 1, i, 0);
    copy_node(d, g.nx, i, 0, i, g.ny - 1);
  } else {
    copy_node(d, g.nx, i, g.ny - 1, i, g.ny - 2);
    copy_node(d, g.nx, i, 0, i, 1);
  }
}

// compute_max_speed (:448-475): the maximum of the nodes' speeds (a NaN speed
// never passes the reference's ">=" test, and fmax ignores it likewise).
// Speeds are >= 0, so their bit patterns order as unsigned integers.  Each
// workgroup reduces its nodes and adds one atomicMax to one of SLOTS slots, a
// cache line apart, so the atomics do not queue on one line; the host takes
// the maximum of the slots.
constexpr int SLOTS = 64, SLOT_STRIDE = 16;     // 16 x 8 B = one 128-byte line

__global__ __launch_bounds__(256) void rd_speed_kernel(Geo g, const double* __restrict__ u,
                                                       unsigned long long* __restrict__ smax) {
  __shared__ double part[4];
  const long long n = (long long)g.nx * g.ny;
  double m = 0.0;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x) {
    double v[4];
    const double4 q = *reinterpret_cast<const double4*>(u + 4 * k);
    v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
    m = fmax(m, speed(g.gamma, v));
  }
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = fmax(fmax(part[0], part[1]), fmax(part[2], part[3]));
    atomicMax(smax + (blockIdx.x % SLOTS) * SLOT_STRIDE, (unsigned long long)__double_as_longlong(m));
  }
}

// dt = min(cfl dx dy/cmax/3, t_end - t) (evolve, :348-349) from the speed
// slots, which it clears for the next step; a maximum that is not finite and
// positive stops the run with the error flag set.
__global__ void rd_dt_kernel(Geo g, double cfl, double t_end, unsigned long long* __restrict__ smax,
                             Clock* __restrict__ ck) {
  if (threadIdx.x != 0) return;
  unsigned long long m = 0;
  for (int k = 0; k < SLOTS; ++k) {
    m = max(m, smax[k * SLOT_STRIDE]);
    smax[k * SLOT_STRIDE] = 0;
  }
  const double cmax = __longlong_as_double((long long)m);
  if (ck->err || !(ck->t < t_end)) {
    ck->active = 0;
    return;
  }
  if (!(cmax > 0.0) || !isfinite(cmax)) {
    ck->err = 1;
    ck->active = 0;
    return;
  }
  ck->dt = fmin(cfl * g.dy * g.dx / cmax / (2.0 * 1.0 + 1.0), t_end - ck->t);
  ck->active = 1;
}
// t = t + dt (:366), after the step
__global__ void rd_tick_kernel(Clock* __restrict__ ck) {
  if (threadIdx.x != 0 || !ck->active) return;
  ck->t = ck->t + ck->dt;
  ck->dt_last = ck->dt;
  ck->steps += 1;
}

// up = u - dudt/(dx dy), diff = up - u (evolve, :351-354)
__global__ void rd_pass1_kernel(long long n, double area, const double* __restrict__ u,
                                const double* __restrict__ dudt, double* __restrict__ up,
                                double* __restrict__ diff, const Clock* __restrict__ ck) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || !ck->active) return;
  const double x = u[k] - dudt[k] / area;
  up[k] = x;
  diff[k] = x - u[k];
}
// u = up - dudt/(dx dy) (:356)
__global__ void rd_pass2_kernel(long long n, double area, double* __restrict__ u, const double* __restrict__ up,
                                const double* __restrict__ dudt, const Clock* __restrict__ ck) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || !ck->active) return;
  u[k] = up[k] - dudt[k] / area;
}

}  // namespace rd

using namespace rd;

struct fvwb_rd {
  fvwb_rd_params p{};
  Geo g{};
  hipStream_t stream = nullptr;
  double* d_u = nullptr;      // state, AoS(4,nx,ny)
  double* d_up = nullptr;
  double* d_diff = nullptr;
  double* d_dudt = nullptr;
  unsigned long long* d_smax = nullptr;
  Clock* d_clk = nullptr;
  bool loaded = false;
  std::string err;
};

static thread_local std::string g_errrd;

static int failrd(fvwb_rd* h, int code, const std::string& msg) {
  (h ? h->err : g_errrd) = msg;
  return code;
}

#define HIPR(h, call)                                                                                  \
  do {                                                                                                 \
    hipError_t e_ = (call);                                                                            \
    if (e_ != hipSuccess) return failrd(h, FVWB_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

extern "C" const char* fvwb_rd_last_error(fvwb_rd_handle h) { return h ? h->err.c_str() : g_errrd.c_str(); }

extern "C" int fvwb_rd_destroy(fvwb_rd_handle h) {
  if (!h) return FVWB_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  void* ps[] = {h->d_u, h->d_up, h->d_diff, h->d_dudt, h->d_smax, h->d_clk};
  for (void* q : ps)
    if (q) (void)hipFree(q);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return FVWB_OK;
}

extern "C" int fvwb_rd_create(fvwb_rd_handle* out, const fvwb_rd_params* p) {
  if (!out || !p) return failrd(nullptr, FVWB_ERR_ARG, "NULL argument");
  *out = nullptr;
  if (p->nx < 2 || p->ny < 2) return failrd(nullptr, FVWB_ERR_ARG, "nx and ny must be >= 2");
  if ((long long)p->nx * p->ny > (1ll << 28)) return failrd(nullptr, FVWB_ERR_ARG, "more than 2^28 nodes");
  if (p->ny > TY * 65535) return failrd(nullptr, FVWB_ERR_ARG, "ny must be <= 458745 (one grid row per 7 node rows)");
  if (p->bc != FVWB_RD_BC_PERIODIC && p->bc != FVWB_RD_BC_TRANSMISSIVE)
    return failrd(nullptr, FVWB_ERR_ARG, "bc must be 2 (periodic copies) or 3 (transmissive copies)");
  if (!(p->gamma > 1.0) || !(p->boxlen_x > 0.0) || !(p->boxlen_y > 0.0) || !(p->cfl > 0.0))
    return failrd(nullptr, FVWB_ERR_ARG, "gamma > 1 and positive box lengths and cfl required");
  fvwb_rd* h = new (std::nothrow) fvwb_rd;
  if (!h) return failrd(nullptr, FVWB_ERR_OOM, "host allocation failed");
  h->p = *p;
  h->g = Geo{p->nx, p->ny, p->bc, p->gamma, p->boxlen_x / (double)(p->nx - 1), p->boxlen_y / (double)(p->ny - 1)};
  const size_t nb = sizeof(double) * 4 * (size_t)p->nx * (size_t)p->ny;
  bool ok = hipStreamCreate(&h->stream) == hipSuccess && hipMalloc(&h->d_u, nb) == hipSuccess &&
            hipMalloc(&h->d_up, nb) == hipSuccess && hipMalloc(&h->d_diff, nb) == hipSuccess &&
            hipMalloc(&h->d_dudt, nb) == hipSuccess &&
            hipMalloc(&h->d_smax, sizeof(unsigned long long) * SLOTS * SLOT_STRIDE) == hipSuccess &&
            hipMalloc(&h->d_clk, sizeof(Clock)) == hipSuccess;
  if (!ok) {
    fvwb_rd_destroy(h);
    return failrd(nullptr, FVWB_ERR_OOM, "device allocation failed");
  }
  *out = h;
  return FVWB_OK;
}

// dudt of compute_update_rd for the device state u and correction diff.
static int update(fvwb_rd* h, const double* u, const double* diff, double dt, const Clock* ck, double* dudt) {
  const dim3 grid((unsigned)((h->p.nx + TX - 1) / TX), (unsigned)((h->p.ny + TY - 1) / TY));
  hipLaunchKernelGGL(rd_update_kernel, grid, dim3(256), 0, h->stream, h->g, u, diff, dt, ck, dudt);
  hipLaunchKernelGGL(rd_bc_x_kernel, dim3((unsigned)((h->p.ny + 255) / 256)), dim3(256), 0, h->stream, h->g, dudt);
  hipLaunchKernelGGL(rd_bc_y_kernel, dim3((unsigned)((h->p.nx + 255) / 256)), dim3(256), 0, h->stream, h->g, dudt);
  HIPR(h, hipGetLastError());
  return FVWB_OK;
}

static void launch_speed(fvwb_rd* h, const double* u) {
  const long long n = (long long)h->p.nx * h->p.ny;
  const unsigned blocks = (unsigned)std::min<long long>((n + 255) / 256, 16384);
  hipLaunchKernelGGL(rd_speed_kernel, dim3(blocks), dim3(256), 0, h->stream, h->g, u, h->d_smax);
}

static int max_speed(fvwb_rd* h, const double* u, double* cmax) {
  const size_t sb = sizeof(unsigned long long) * SLOTS * SLOT_STRIDE;
  HIPR(h, hipMemsetAsync(h->d_smax, 0, sb, h->stream));
  launch_speed(h, u);
  HIPR(h, hipGetLastError());
  unsigned long long bits[SLOTS * SLOT_STRIDE];
  HIPR(h, hipMemcpyAsync(bits, h->d_smax, sb, hipMemcpyDeviceToHost, h->stream));
  HIPR(h, hipStreamSynchronize(h->stream));
  unsigned long long m = 0;
  for (int k = 0; k < SLOTS; ++k) m = std::max(m, bits[k * SLOT_STRIDE]);
  std::memcpy(cmax, &m, sizeof m);
  return FVWB_OK;
}

extern "C" int fvwb_rd_compute_update(fvwb_rd_handle h, const double* u, const double* diff, double dt,
                                      double* dudt) {
  if (!h || !u || !diff || !dudt) return failrd(h, FVWB_ERR_ARG, "NULL argument");
  const size_t nb = sizeof(double) * 4 * (size_t)h->p.nx * (size_t)h->p.ny;
  // the update works in the step's scratch arrays; an uploaded state stays as it is
  HIPR(h, hipMemcpyAsync(h->d_up, u, nb, hipMemcpyHostToDevice, h->stream));
  HIPR(h, hipMemcpyAsync(h->d_diff, diff, nb, hipMemcpyHostToDevice, h->stream));
  int rc = update(h, h->d_up, h->d_diff, dt, nullptr, h->d_dudt);
  if (rc) return rc;
  HIPR(h, hipMemcpyAsync(dudt, h->d_dudt, nb, hipMemcpyDeviceToHost, h->stream));
  HIPR(h, hipStreamSynchronize(h->stream));
  return FVWB_OK;
}

extern "C" int fvwb_rd_compute_max_speed(fvwb_rd_handle h, const double* u, double* cmax) {
  if (!h || !u || !cmax) return failrd(h, FVWB_ERR_ARG, "NULL argument");
  const size_t nb = sizeof(double) * 4 * (size_t)h->p.nx * (size_t)h->p.ny;
  HIPR(h, hipMemcpyAsync(h->d_up, u, nb, hipMemcpyHostToDevice, h->stream));
  return max_speed(h, h->d_up, cmax);
}

extern "C" int fvwb_rd_upload(fvwb_rd_handle h, const double* u) {
  if (!h || !u) return failrd(h, FVWB_ERR_ARG, "NULL argument");
  const size_t nb = sizeof(double) * 4 * (size_t)h->p.nx * (size_t)h->p.ny;
  const Clock zero{};
  HIPR(h, hipMemcpyAsync(h->d_u, u, nb, hipMemcpyHostToDevice, h->stream));
  HIPR(h, hipMemcpyAsync(h->d_clk, &zero, sizeof zero, hipMemcpyHostToDevice, h->stream));
  HIPR(h, hipStreamSynchronize(h->stream));
  h->loaded = true;
  return FVWB_OK;
}

static int read_clock(fvwb_rd* h, Clock* c) {
  HIPR(h, hipMemcpyAsync(c, h->d_clk, sizeof *c, hipMemcpyDeviceToHost, h->stream));
  HIPR(h, hipStreamSynchronize(h->stream));
  return FVWB_OK;
}

// Steps are queued in groups of this many between reads of the clock.
static constexpr int32_t kGroup = 64;

// evolve (:356-385) with solver 'RD1': while t < t_end, at most nsteps steps.
// dt, t and the step count live on the device; the host reads them once per
// group of steps to stop early, and each kernel of a step past t_end (or
// after a failure) does nothing.
extern "C" int fvwb_rd_advance(fvwb_rd_handle h, int32_t nsteps, double t_end) {
  if (!h) return failrd(h, FVWB_ERR_ARG, "NULL handle");
  if (!h->loaded) return failrd(h, FVWB_ERR_STATE, "advance before upload");
  if (nsteps < 0) return failrd(h, FVWB_ERR_ARG, "nsteps must be >= 0");
  const long long n = 4ll * h->p.nx * h->p.ny;
  const double area = h->g.dx * h->g.dy;
  const dim3 eb(256), eg((unsigned)((n + 255) / 256));
  HIPR(h, hipMemsetAsync(h->d_smax, 0, sizeof(unsigned long long) * SLOTS * SLOT_STRIDE, h->stream));
  Clock c{};
  int rc = read_clock(h, &c);
  if (rc) return rc;
  int32_t done = 0;
  while (done < nsteps && !c.err && c.t < t_end) {
    const int32_t k = std::min(kGroup, nsteps - done);
    for (int32_t s = 0; s < k; ++s) {
      launch_speed(h, h->d_u);
      hipLaunchKernelGGL(rd_dt_kernel, dim3(1), dim3(64), 0, h->stream, h->g, h->p.cfl, t_end, h->d_smax, h->d_clk);
      HIPR(h, hipMemsetAsync(h->d_diff, 0, sizeof(double) * (size_t)n, h->stream));
      if ((rc = update(h, h->d_u, h->d_diff, 0.0, h->d_clk, h->d_dudt))) return rc;
      hipLaunchKernelGGL(rd_pass1_kernel, eg, eb, 0, h->stream, n, area, h->d_u, h->d_dudt, h->d_up, h->d_diff,
                         h->d_clk);
      if ((rc = update(h, h->d_u, h->d_diff, 0.0, h->d_clk, h->d_dudt))) return rc;
      hipLaunchKernelGGL(rd_pass2_kernel, eg, eb, 0, h->stream, n, area, h->d_u, h->d_up, h->d_dudt, h->d_clk);
      hipLaunchKernelGGL(rd_tick_kernel, dim3(1), dim3(64), 0, h->stream, h->d_clk);
      HIPR(h, hipGetLastError());
    }
    done += k;
    if ((rc = read_clock(h, &c))) return rc;
  }
  if (c.err) return failrd(h, FVWB_ERR_NONPHYSICAL, "wave-speed maximum not finite and positive");
  return FVWB_OK;
}

extern "C" int fvwb_rd_download(fvwb_rd_handle h, double* u) {
  if (!h || !u) return failrd(h, FVWB_ERR_ARG, "NULL argument");
  if (!h->loaded) return failrd(h, FVWB_ERR_STATE, "download before upload");
  const size_t nb = sizeof(double) * 4 * (size_t)h->p.nx * (size_t)h->p.ny;
  HIPR(h, hipMemcpyAsync(u, h->d_u, nb, hipMemcpyDeviceToHost, h->stream));
  HIPR(h, hipStreamSynchronize(h->stream));
  return FVWB_OK;
}

// ---- checkpoint / resume: the state and the clock -----------------------------------
extern "C" int fvwb_rd_checkpoint_save(fvwb_rd_handle h, const char* path) {
  if (!h || !path) return failrd(h, FVWB_ERR_ARG, "NULL argument");
  if (!h->loaded) return failrd(h, FVWB_ERR_STATE, "checkpoint before upload");
  const fvwb_ckpt::Segment seg[2] = {{h->d_u, sizeof(double) * 4 * (size_t)h->p.nx * (size_t)h->p.ny},
                                     {h->d_clk, sizeof(Clock)}};
  std::string err;
  const int rc = fvwb_ckpt::save(path, fvwb_ckpt::RD, &h->p, sizeof h->p, 0, seg, 2, h->stream, err);
  if (rc) h->err = err;
  return rc;
}

extern "C" int fvwb_rd_checkpoint_load(fvwb_rd_handle h, const char* path) {
  if (!h || !path) return failrd(h, FVWB_ERR_ARG, "NULL argument");
  const fvwb_ckpt::Segment seg[2] = {{h->d_u, sizeof(double) * 4 * (size_t)h->p.nx * (size_t)h->p.ny},
                                     {h->d_clk, sizeof(Clock)}};
  std::string err;
  int32_t parity = 0;
  const int rc = fvwb_ckpt::load(path, fvwb_ckpt::RD, &h->p, sizeof h->p, &parity, seg, 2, h->stream, err);
  h->loaded = rc == FVWB_OK;
  if (rc) h->err = err;
  return rc;
}

extern "C" int fvwb_rd_get_time(fvwb_rd_handle h, double* t, double* dt_last, int64_t* steps) {
  if (!h) return failrd(h, FVWB_ERR_ARG, "NULL handle");
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
