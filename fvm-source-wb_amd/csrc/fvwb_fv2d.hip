// fvwb_fv2d.hip — 2D finite-volume kernels (gfx950, wave64).
//
// One fused kernel per Runge-Kutta stage computes, for every cell, the
// well-balanced residual of compute_update_exact (benchmark_2d.f90:465-618) —
// reconstruction of the deviation dU at faces, numerical flux minus the
// equilibrium flux, source deviation — and the stage combination
// out = alpha*dU^n + beta*(dU_in + dt*L(dU_in)).  The last stage also reduces
// max(|v|+c) of the new state for the next dt (compute_max_speed,
// benchmark_2d.f90:264-279), so a whole run is enqueued without host syncs.
//
// Work decomposition: a wave owns a strip of 64 columns (2H of them halo,
// H = 1 first order, 2 MUSCL) and `rows` rows, and marches upward with a
// sliding register window.  x-neighbours come from DPP lane moves; y-neighbours
// from the window; each row is one coalesced 64-wide load per variable plane.
// Face fluxes are computed once: the x face on a lane's left, the y face above
// the current row (carried to the next row).
#include "fvwb_device.h"
#include "fvwb_internal.h"

namespace fvwb {

enum { BC_FROZEN = 0, BC_REFLECT = 1 };
enum { MODE_RESIDUAL = 0, MODE_STAGE1 = 1, MODE_STAGE = 2 };

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Blocks are dealt round-robin to the 8 XCDs; renumber them so that each XCD
// receives a contiguous range of strips (neighbouring halos share its L2).
__device__ __forceinline__ int xcd_block(int b, int nb) {
  int xcd = b & 7, idx = b >> 3, base = nb >> 3, rem = nb & 7;
  return xcd * base + (xcd < rem ? xcd : rem) + idx;
}

// Row access through a buffer descriptor built on the scalar unit: base = the
// row's first element (uniform), extent = the row, per-lane 32-bit byte offset.
// No vector address arithmetic; a lane offset past the row would read 0 / be
// dropped by the hardware range check.
template <typename T>
__device__ __forceinline__ T row_load(const T* row, unsigned boff, int row_bytes) {
  __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc((void*)row, 0, row_bytes, 0x00020000);
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, boff, 0, 0));
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, boff, 0, 0));
  }
}
template <typename T>
__device__ __forceinline__ void row_store(T* row, unsigned boff, int row_bytes, T v) {
  __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc((void*)row, 0, row_bytes, 0x00020000);
  if constexpr (sizeof(T) == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, boff, 0, 0);
  } else {
    typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, boff, 0, 0);
  }
}

template <typename T, int RECON, int FLUX, int BC, int FORM, int MODE>
__global__ __launch_bounds__(256) void fv2d_stage_kernel(StageArgs<T> a) {
  constexpr int H = RECON ? 2 : 1;
  constexpr int COLS = 64 - 2 * H;
  const int lane = threadIdx.x & 63;
  // wave index made explicitly uniform so that strip, row and row pointers
  // live in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

  // ---- step control (dt from the previous step's reduction) -----------------
  double dt = 0.0;
  const int p = a.parity, q = p ^ 1;
  if (MODE != MODE_RESIDUAL) {
    const Ctrl* c = a.ctrl;
    const double t = c->t[p];
    const bool active = t < a.t_end;
    unsigned long long m = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) m = c->cmax[p][k] > m ? c->cmax[p][k] : m;
    // dt = 0.5*dx/cmax*cfl (benchmark_2d.f90:242)
    dt = 0.5 * a.h / __longlong_as_double((long long)m) * a.cfl;
    if (blockIdx.x == 0 && threadIdx.x < 8) {
      Ctrl* cw = a.ctrl;
      if (a.first) cw->cmax[q][threadIdx.x] = 0ull;
      if (a.last && threadIdx.x == 0) {
        cw->t[q] = active ? t + dt : t;
        if (active) { cw->steps_done += 1; cw->dt_last = dt; }
      }
    }
    if (!active) return;
  }

  const int tile = xcd_block(blockIdx.x, gridDim.x) * 4 + wave;
  if (tile >= a.ntiles) return;
  const int tx = tile % a.ntx, ty = tile / a.ntx;
  const int nx = a.nx, ny = a.ny;
  const int i = tx * COLS + lane - H;          // this lane's column
  const int y0 = ty * a.rows;
  const int y1 = min(y0 + a.rows, ny);
  const bool own = lane >= H && lane < 64 - H && i < nx;

  // column with the boundary treatment (mirror for walls, clamp otherwise)
  int icol;
  T sx = T(1);
  if (BC == BC_REFLECT) {
    icol = i < 0 ? -1 - i : (i >= nx ? 2 * nx - 1 - i : i);
    if (i < 0 || i >= nx) sx = T(-1);
    icol = clampi(icol, 0, nx - 1);
  } else {
    icol = clampi(i, 0, nx - 1);
  }
  const unsigned boff = (unsigned)icol * (unsigned)sizeof(T);
  const int rowb = nx * (int)sizeof(T);
  const long long plane = a.plane;
  const int pitch = a.pitch;

  auto load_row = [&](const T* base, int j) -> Vec4<T> {
    int jc;
    bool flip = false;
    if (BC == BC_REFLECT) {
      jc = j < 0 ? -1 - j : (j >= ny ? 2 * ny - 1 - j : j);
      flip = j < 0 || j >= ny;
      jc = clampi(jc, 0, ny - 1);
    } else {
      jc = clampi(j, 0, ny - 1);
    }
    const T* r = base + (long long)jc * pitch;
    Vec4<T> v = {row_load(r, boff, rowb), row_load(r + plane, boff, rowb),
                 row_load(r + 2 * plane, boff, rowb), row_load(r + 3 * plane, boff, rowb)};
    if (BC == BC_REFLECT) {
      v.mx *= sx;
      if (flip) v.my = -v.my;
    }
    return v;
  };
  // cell-centred rows (clamped into the grid; rows past the strip are masked)
  auto load_cell = [&](const T* base, int j) -> Vec4<T> {
    const T* r = base + (long long)min(j, ny - 1) * pitch;
    return Vec4<T>{row_load(r, boff, rowb), row_load(r + plane, boff, rowb),
                   row_load(r + 2 * plane, boff, rowb), row_load(r + 3 * plane, boff, rowb)};
  };

  Phys<T> P;
  P.gamma = T(a.gamma);
  P.gm1 = T(a.gamma - 1.0);
  P.half = T(0.5);
  P.g = T(a.g);
  const T odx = T(a.odx), ody = T(a.ody);
  // stage combination out = alpha*U^n + beta*D + (beta*dt)*L
  const T alpha = T(a.alpha), beta = T(a.beta), bdt = T(a.beta * dt);

  // separable FP64 equilibrium: rho = (rho0 e^{-kx}) e^{-ky} is formed in FP64
  // and rounded once; p, E and E+p scale with rho by constant ratios.
  const double rxf = a.rxf[clampi(i, 0, nx)];       // left x face of this column
  const double rxc = a.rxc[clampi(i, 0, nx - 1)];
  const T por = T(a.p_over_rho), eor = T(a.p_over_rho * a.inv_gm1);
  const T epor = T(a.p_over_rho * a.gamma * a.inv_gm1);
  auto face_eq = [&](double rx, double ey) -> FaceEq<T> {
    FaceEq<T> f;
    f.rho = T(rx * ey);
    f.p = f.rho * por;
    f.e = f.rho * eor;
    f.ep = f.rho * epor;
    return f;
  };

  double smax = 0.0;
  bool bad = false;

  // ---- prologue -----------------------------------------------------------------
  // A ring of RING rows holds rows j..j+RING-2 while row j is computed and
  // receives row j+RING-1 (so RING-3 rows are in flight ahead of use); U^n uses
  // a ring of URING rows the same way.  The body is unrolled RING times with
  // rotated roles, so the rings never move data between registers.
  constexpr int RING = sizeof(T) == 4 ? 6 : 4;
  constexpr int URING = sizeof(T) == 4 ? 3 : 2;
  Vec4<T> R[RING], Ur[URING], sy, Gp, dfw;
  {
    FaceEq<T> qy = face_eq(rxc, a.eyf[clampi(y0, 0, ny)]);
    Vec4<T> Dm1 = load_row(a.in, y0 - 1);
#pragma unroll
    for (int k = 0; k < RING - 1; ++k) R[k] = load_row(a.in, y0 + k);
    if (RECON) {
      Vec4<T> Dm2 = load_row(a.in, y0 - 2);
      Vec4<T> sym1 = minmod2(Dm1 - Dm2, R[0] - Dm1);
      sy = minmod2(R[0] - Dm1, R[1] - R[0]);
      dfw = R[1] - R[0];   // B - A of the first row
      Gp = face_flux<T, FLUX, 1, FORM>(P, qy, Dm1 + scale(P.half, sym1), R[0] - scale(P.half, sy));
    } else {
      Gp = face_flux<T, FLUX, 1, FORM>(P, qy, Dm1, R[0]);
    }
#pragma unroll
    for (int k = 0; k < URING; ++k) Ur[k] = Vec4<T>{T(0), T(0), T(0), T(0)};
    if (MODE == MODE_STAGE) {
#pragma unroll
      for (int k = 0; k < URING - 1; ++k) Ur[k] = load_cell(a.un, y0 + k);
    }
  }

  // one row: A = row j, B = row j+1, C = row j+2; S receives row j+RING-1; UA
  // holds U^n(j) and US receives U^n(j+URING-1).
  auto row_step = [&](int j, const Vec4<T>& A, const Vec4<T>& B, const Vec4<T>& C, Vec4<T>& S,
                      const Vec4<T>& UA, Vec4<T>& US) {
    S = load_row(a.in, j + RING - 1);
    if (MODE == MODE_STAGE) US = load_cell(a.un, j + URING - 1);
    const int jt = min(j, ny - 1);
    const double eyc = a.eyc[jt];
    const double eyf = a.eyf[jt + 1];

    // y face above row j
    Vec4<T> Gn, xl, xr;
    if (RECON) {
      Vec4<T> dnext = C - B;
      Vec4<T> syn = minmod2(dfw, dnext);   // dfw = B - A, carried from the previous row
      dfw = dnext;
      Gn = face_flux<T, FLUX, 1, FORM>(P, face_eq(rxc, eyf), A + scale(P.half, sy), B - scale(P.half, syn));
      sy = syn;
      // x slopes on row j from the neighbouring lanes
      Vec4<T> sxl = minmod2(A - from_left(A), from_right(A) - A);
      xr = A + scale(P.half, sxl);   // state at this cell's right face
      xl = A - scale(P.half, sxl);   // state at this cell's left face
    } else {
      Gn = face_flux<T, FLUX, 1, FORM>(P, face_eq(rxc, eyf), A, B);
      xr = A;
      xl = A;
    }
    // x face on this lane's left: left state from the neighbour's right face
    Vec4<T> Fl = face_flux<T, FLUX, 0, FORM>(P, face_eq(rxf, eyc), from_left(xr), xl);
    Vec4<T> Fr = from_right(Fl);

    // residual: -dF/dx - dG/dy + (s(U) - s(U_eq)), s = (0, -rho g, -rho g, -g (m_x+m_y))
    Vec4<T> L;
    L.r = -(Fr.r - Fl.r) * odx - (Gn.r - Gp.r) * ody;
    L.mx = (-(Fr.mx - Fl.mx) * odx - (Gn.mx - Gp.mx) * ody) - P.g * A.r;
    L.my = (-(Fr.my - Fl.my) * odx - (Gn.my - Gp.my) * ody) - P.g * A.r;
    L.e = (-(Fr.e - Fl.e) * odx - (Gn.e - Gp.e) * ody) - P.g * (A.mx + A.my);
    Gp = Gn;
    if (BC == BC_FROZEN && (i == 0 || i == nx - 1 || j == 0 || j == ny - 1)) L = {T(0), T(0), T(0), T(0)};

    Vec4<T> o;
    if (MODE == MODE_RESIDUAL) {
      o = L;
    } else if (MODE == MODE_STAGE1) {
      o = A + scale(bdt, L);
    } else {
      o = scale(alpha, UA) + (scale(beta, A) + scale(bdt, L));
    }
    if (own && j < y1) {
      T* d = a.out + (long long)j * pitch;
      row_store(d, boff, rowb, o.r);
      row_store(d + plane, boff, rowb, o.mx);
      row_store(d + 2 * plane, boff, rowb, o.my);
      row_store(d + 3 * plane, boff, rowb, o.e);
      if (MODE != MODE_RESIDUAL && a.last) {
        bool ok;
        double rc = rxc * eyc;
        T rct = T(rc);
        T sp = cell_speed(P, rct, rct * eor, o, ok);
        smax = fmax(smax, (double)sp);
        bad |= !ok;
      }
    }
  };

  for (int j = y0; j < y1; j += RING) {
#pragma unroll
    for (int k = 0; k < RING; ++k)
      row_step(j + k, R[k % RING], R[(k + 1) % RING], R[(k + 2) % RING], R[(k + RING - 1) % RING],
               Ur[k % URING], Ur[(k + URING - 1) % URING]);
  }

  if (MODE != MODE_RESIDUAL && a.last) {
    if (bad) atomicOr(&a.ctrl->err, 1u);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) smax = fmax(smax, __shfl_xor(smax, off));
    if (lane == 0) atomicMax(&a.ctrl->cmax[q][blockIdx.x & 7], (unsigned long long)__double_as_longlong(smax));
  }
}

// ---- auxiliary kernels ---------------------------------------------------------

// dU = u - conservative(w_eq) (benchmark_2d.f90:496-499), FP64, into planes of T.
template <typename T>
__global__ void aos_to_dev_kernel(const double* __restrict__ u, const double* __restrict__ w,
                                  T* __restrict__ out, int nx, int ny, int pitch, long long plane,
                                  double gamma) {
  long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (long long)nx * ny) return;
  int i = (int)(c % nx), j = (int)(c / nx);
  const double* uc = u + 4 * c;
  const double* wc = w + 4 * c;
  double ue0 = wc[0], ue1 = wc[0] * wc[1], ue2 = wc[0] * wc[2];
  double ue3 = wc[3] / (gamma - 1.0) + 0.5 * (wc[0] * (wc[1] * wc[1] + wc[2] * wc[2]));
  long long o = (long long)j * pitch + i;
  out[o] = T(uc[0] - ue0);
  out[o + plane] = T(uc[1] - ue1);
  out[o + 2 * plane] = T(uc[2] - ue2);
  out[o + 3 * plane] = T(uc[3] - ue3);
}

// AoS(4,nx,ny) FP64 = [conservative(w_eq) +] planes (widened exactly).
template <typename T, bool ADD_EQ>
__global__ void dev_to_aos_kernel(const T* __restrict__ in, const double* __restrict__ w,
                                  double* __restrict__ u, int nx, int ny, int pitch, long long plane,
                                  double gamma) {
  long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (long long)nx * ny) return;
  int i = (int)(c % nx), j = (int)(c / nx);
  long long o = (long long)j * pitch + i;
  double d0 = (double)in[o], d1 = (double)in[o + plane], d2 = (double)in[o + 2 * plane],
         d3 = (double)in[o + 3 * plane];
  if (ADD_EQ) {
    const double* wc = w + 4 * c;
    double ue0 = wc[0], ue1 = wc[0] * wc[1], ue2 = wc[0] * wc[2];
    double ue3 = wc[3] / (gamma - 1.0) + 0.5 * (wc[0] * (wc[1] * wc[1] + wc[2] * wc[2]));
    d0 += ue0; d1 += ue1; d2 += ue2; d3 += ue3;
  }
  double* uc = u + 4 * c;
  uc[0] = d0; uc[1] = d1; uc[2] = d2; uc[3] = d3;
}

// compute_max_speed (benchmark_2d.f90:264-279) over an AoS FP64 state.
__global__ void max_speed_aos_kernel(const double* __restrict__ u, long long n, double gamma,
                                     unsigned long long* __restrict__ out) {
  double m = 0.0;
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (long long)gridDim.x * blockDim.x) {
    const double* s = u + 4 * c;
    double w0 = s[0], w1 = s[1] / w0, w2 = s[2] / w0;
    double w3 = (gamma - 1.0) * (s[3] - 0.5 * w0 * (w1 * w1 + w2 * w2));
    double cs = sqrt(gamma * fmax(w3, 1e-10) / fmax(w0, 1e-10));
    m = fmax(m, sqrt(w1 * w1 + w2 * w2) + cs);
  }
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

// Initial reduction of max(|v|+c) of the stored state into ctrl->cmax[0].
template <typename T>
__global__ void cfl_init_kernel(const T* __restrict__ in, const double* __restrict__ rxc,
                                const double* __restrict__ eyc, Ctrl* ctrl, int nx, int ny,
                                int pitch, long long plane, double gamma, double p_over_rho,
                                double inv_gm1) {
  Phys<T> P;
  P.gamma = T(gamma); P.gm1 = T(gamma - 1.0); P.half = T(0.5); P.g = T(0);
  double m = 0.0;
  bool bad = false;
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < (long long)nx * ny;
       c += (long long)gridDim.x * blockDim.x) {
    int i = (int)(c % nx), j = (int)(c / nx);
    long long o = (long long)j * pitch + i;
    Vec4<T> d = {in[o], in[o + plane], in[o + 2 * plane], in[o + 3 * plane]};
    double rc = rxc[i] * eyc[j];
    bool ok;
    T sp = cell_speed(P, T(rc), T(rc * p_over_rho * inv_gm1), d, ok);
    m = fmax(m, (double)sp);
    bad |= !ok;
  }
  if (bad) atomicOr(&ctrl->err, 1u);
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0)
    atomicMax(&ctrl->cmax[0][blockIdx.x & 7], (unsigned long long)__double_as_longlong(m));
}

// Device-side initial condition: w_eq = closed-form equilibrium at centres
// (get_equilibrium_solution, benchmark_2d.f90:195-202) and dU = (0,0,0,dp/(gamma-1))
// with dp = amp*exp(-width*((x-x0)^2+(y-y0)^2)).
template <typename T>
__global__ void init_gaussian_kernel(T* __restrict__ out, double* __restrict__ w, int nx, int ny,
                                     int pitch, long long plane, double dx, double dy, double rho0,
                                     double p0, double kk, double gamma, double amp, double x0,
                                     double y0, double width) {
  long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (long long)nx * ny) return;
  int i = (int)(c % nx), j = (int)(c / nx);
  double x = ((double)i + 0.5) * dx, y = ((double)j + 0.5) * dy;
  double e = exp(-(kk * (x + y)));
  double* wc = w + 4 * c;
  wc[0] = rho0 * e; wc[1] = 0.0; wc[2] = 0.0; wc[3] = p0 * e;
  double dp = amp * exp(-width * ((x - x0) * (x - x0) + (y - y0) * (y - y0)));
  long long o = (long long)j * pitch + i;
  out[o] = T(0); out[o + plane] = T(0); out[o + 2 * plane] = T(0);
  out[o + 3 * plane] = T(dp / (gamma - 1.0));
}

// ---- host-side launchers ----------------------------------------------------------

template <typename T, int RECON, int FLUX, int BC, int FORM, int MODE>
static hipError_t launch_stage_t(const StageArgs<T>& a, hipStream_t s) {
  int blocks = (a.ntiles + 3) / 4;
  hipLaunchKernelGGL((fv2d_stage_kernel<T, RECON, FLUX, BC, FORM, MODE>), dim3(blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

template <typename T, int FORM, int MODE>
static hipError_t dispatch_scheme(const StageArgs<T>& a, int recon, int flux, int bc, hipStream_t s) {
#define FVWB_CASE(R, F, B)                                              \
  if (recon == R && flux == F && bc == B) return launch_stage_t<T, R, F, B, FORM, MODE>(a, s);
  FVWB_CASE(0, 0, 0) FVWB_CASE(0, 0, 1) FVWB_CASE(0, 1, 0) FVWB_CASE(0, 1, 1)
  FVWB_CASE(1, 0, 0) FVWB_CASE(1, 0, 1) FVWB_CASE(1, 1, 0) FVWB_CASE(1, 1, 1)
#undef FVWB_CASE
  return hipErrorInvalidValue;
}

// Stage modes: 0 residual (FP64 plain only), 1 first stage, 2 later stages.
hipError_t launch_fv2d_stage_f64(const StageArgs<double>& a, int recon, int flux, int bc, int mode,
                                 hipStream_t s) {
  if (mode == MODE_RESIDUAL) return dispatch_scheme<double, FORM_PLAIN, MODE_RESIDUAL>(a, recon, flux, bc, s);
  if (mode == MODE_STAGE1) return dispatch_scheme<double, FORM_PLAIN, MODE_STAGE1>(a, recon, flux, bc, s);
  return dispatch_scheme<double, FORM_PLAIN, MODE_STAGE>(a, recon, flux, bc, s);
}

hipError_t launch_fv2d_stage_f32(const StageArgs<float>& a, int recon, int flux, int bc, int mode,
                                 hipStream_t s) {
  if (mode == MODE_STAGE1) return dispatch_scheme<float, FORM_DEV, MODE_STAGE1>(a, recon, flux, bc, s);
  if (mode == MODE_STAGE) return dispatch_scheme<float, FORM_DEV, MODE_STAGE>(a, recon, flux, bc, s);
  return hipErrorInvalidValue;
}

static inline unsigned grid1d(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <typename T>
hipError_t launch_aos_to_dev(const double* u, const double* w, T* out, int nx, int ny, int pitch,
                             long long plane, double gamma, hipStream_t s) {
  long long n = (long long)nx * ny;
  hipLaunchKernelGGL((aos_to_dev_kernel<T>), dim3(grid1d(n, 256)), dim3(256), 0, s, u, w, out, nx, ny,
                     pitch, plane, gamma);
  return hipGetLastError();
}
template hipError_t launch_aos_to_dev<double>(const double*, const double*, double*, int, int, int, long long, double, hipStream_t);
template hipError_t launch_aos_to_dev<float>(const double*, const double*, float*, int, int, int, long long, double, hipStream_t);

template <typename T>
hipError_t launch_dev_to_aos(const T* in, const double* w, double* u, int nx, int ny, int pitch,
                             long long plane, double gamma, bool add_eq, hipStream_t s) {
  long long n = (long long)nx * ny;
  if (add_eq)
    hipLaunchKernelGGL((dev_to_aos_kernel<T, true>), dim3(grid1d(n, 256)), dim3(256), 0, s, in, w, u, nx,
                       ny, pitch, plane, gamma);
  else
    hipLaunchKernelGGL((dev_to_aos_kernel<T, false>), dim3(grid1d(n, 256)), dim3(256), 0, s, in, w, u, nx,
                       ny, pitch, plane, gamma);
  return hipGetLastError();
}
template hipError_t launch_dev_to_aos<double>(const double*, const double*, double*, int, int, int, long long, double, bool, hipStream_t);
template hipError_t launch_dev_to_aos<float>(const float*, const double*, double*, int, int, int, long long, double, bool, hipStream_t);

hipError_t launch_max_speed_aos(const double* u, long long n, double gamma, unsigned long long* out,
                                hipStream_t s) {
  unsigned blocks = grid1d(n, 256);
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(max_speed_aos_kernel, dim3(blocks), dim3(256), 0, s, u, n, gamma, out);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_cfl_init(const T* in, const double* rxc, const double* eyc, Ctrl* ctrl, int nx, int ny,
                           int pitch, long long plane, double gamma, double p_over_rho, double inv_gm1,
                           hipStream_t s) {
  unsigned blocks = grid1d((long long)nx * ny, 256);
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL((cfl_init_kernel<T>), dim3(blocks), dim3(256), 0, s, in, rxc, eyc, ctrl, nx, ny, pitch,
                     plane, gamma, p_over_rho, inv_gm1);
  return hipGetLastError();
}
template hipError_t launch_cfl_init<double>(const double*, const double*, const double*, Ctrl*, int, int, int, long long, double, double, double, hipStream_t);
template hipError_t launch_cfl_init<float>(const float*, const double*, const double*, Ctrl*, int, int, int, long long, double, double, double, hipStream_t);

template <typename T>
hipError_t launch_init_gaussian(T* out, double* w, int nx, int ny, int pitch, long long plane, double dx,
                                double dy, double rho0, double p0, double kk, double gamma, double amp,
                                double x0, double y0, double width, hipStream_t s) {
  long long n = (long long)nx * ny;
  hipLaunchKernelGGL((init_gaussian_kernel<T>), dim3(grid1d(n, 256)), dim3(256), 0, s, out, w, nx, ny, pitch,
                     plane, dx, dy, rho0, p0, kk, gamma, amp, x0, y0, width);
  return hipGetLastError();
}
template hipError_t launch_init_gaussian<double>(double*, double*, int, int, int, long long, double, double, double, double, double, double, double, double, double, double, hipStream_t);
template hipError_t launch_init_gaussian<float>(float*, double*, int, int, int, long long, double, double, double, double, double, double, double, double, double, double, hipStream_t);

}  // namespace fvwb

namespace fvwb {

// Plain streaming copy (16 B per lane per iteration): the calibration point for
// the achieved-bandwidth and HBM-counter figures of the solver kernels.
__global__ __launch_bounds__(256) void copy_probe_kernel(const float4* __restrict__ src, float4* __restrict__ dst,
                                                         long long n4) {
  // each thread moves 4 float4 per iteration, all four loads issued before the
  // stores, so a wave keeps 4 KiB of reads in flight
  const long long stride = (long long)gridDim.x * blockDim.x;
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  for (; k + 3 * stride < n4; k += 4 * stride) {
    float4 a = src[k], b = src[k + stride], c = src[k + 2 * stride], d = src[k + 3 * stride];
    dst[k] = a;
    dst[k + stride] = b;
    dst[k + 2 * stride] = c;
    dst[k + 3 * stride] = d;
  }
  for (; k < n4; k += stride) dst[k] = src[k];
}

hipError_t launch_copy_probe(const void* src, void* dst, long long bytes, hipStream_t s) {
  long long n4 = bytes / 16;
  hipLaunchKernelGGL(copy_probe_kernel, dim3(256 * 16), dim3(256), 0, s, (const float4*)src, (float4*)dst, n4);
  return hipGetLastError();
}

}  // namespace fvwb
