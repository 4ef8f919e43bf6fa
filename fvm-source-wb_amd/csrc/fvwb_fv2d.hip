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
#include <type_traits>

#include "fvwb_device.h"
#include "fvwb_internal.h"

namespace fvwb {

enum { BC_FROZEN = 0, BC_REFLECT = 1 };
enum { MODE_RESIDUAL = 0, MODE_STAGE1 = 1, MODE_STAGE = 2 };

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Read of a table that no kernel writes while a stage runs, through the
// constant address space: at a wave-uniform index it becomes a scalar load,
// which waits on its own counter and so never holds up on the row loads in
// flight (a vector load would, as vector loads complete in order).
__device__ __forceinline__ double table_at(const double* t, int k) {
  return ((const __attribute__((address_space(4))) double*)t)[k];
}

// Blocks are dealt round-robin to the 8 XCDs; renumber them so that each XCD
// receives a contiguous range of strips (neighbouring halos share its L2).
__device__ __forceinline__ int xcd_block(int b, int nb) {
  int xcd = b & 7, idx = b >> 3, base = nb >> 3, rem = nb & 7;
  return xcd * base + (xcd < rem ? xcd : rem) + idx;
}

// Row access through a buffer descriptor built on the scalar unit: base = the
// row's first element (uniform), extent = one variable's row, per-lane
// 32-bit byte offset.  No vector address arithmetic, and the hardware range
// check bounds every access to the row.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const void* row, int row_bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)row, 0, row_bytes, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T ld1(const T* row, unsigned boff, int row_bytes) {
  if constexpr (sizeof(T) == 4)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(row_rsrc(row, row_bytes), boff, 0, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(row_rsrc(row, row_bytes), boff, 0, 0));
}

// A lane's block of C consecutive columns of one variable: one 4-, 8- or 16-byte load.
template <typename T, int C>
struct Lanes { T v[C]; };
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
// AUX is the buffer instruction's cache policy (0 = default).  Nontemporal
// policies were measured and did not help the stage kernels.
template <typename T, int C, int AUX = 0>
__device__ __forceinline__ Lanes<T, C> ldv(const T* row, unsigned boff, int row_bytes) {
  static_assert(C * sizeof(T) == 4 || C * sizeof(T) == 8 || C * sizeof(T) == 16, "block of 4, 8 or 16 bytes");
  if constexpr (C * sizeof(T) == 16)
    return __builtin_bit_cast(Lanes<T, C>, __builtin_amdgcn_raw_buffer_load_b128(row_rsrc(row, row_bytes), boff, 0, AUX));
  else if constexpr (C * sizeof(T) == 8)
    return __builtin_bit_cast(Lanes<T, C>, __builtin_amdgcn_raw_buffer_load_b64(row_rsrc(row, row_bytes), boff, 0, AUX));
  else
    return __builtin_bit_cast(Lanes<T, C>, __builtin_amdgcn_raw_buffer_load_b32(row_rsrc(row, row_bytes), boff, 0, AUX));
}
template <typename T, int C, int AUX = 0>
__device__ __forceinline__ void stv(T* row, unsigned boff, int row_bytes, Lanes<T, C> x) {
  if constexpr (C * sizeof(T) == 16)
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, x), row_rsrc(row, row_bytes), boff, 0, AUX);
  else if constexpr (C * sizeof(T) == 8)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), row_rsrc(row, row_bytes), boff, 0, AUX);
  else
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), row_rsrc(row, row_bytes), boff, 0, AUX);
}

// Deviation state of a lane's C columns on one row.
template <typename T, int C>
struct Blk { Vec4<T> c[C]; };

// One wave computes one strip ("tile") of one instance.  The wave that owns
// tile 0 of an instance also does that instance's step bookkeeping: the first
// stage clears the next step's speed slots, the last stage advances t.  LAST
// (the step's last stage, a.last) is a template parameter so that the other
// stages carry no wave-speed code at all.
template <typename T, int RECON, int FLUX, int BC, int FORM, int MODE, int CPL, bool LAST>
__device__ __forceinline__ void stage_tile(const StageArgs<T>& a, const int tile, const int lane, const int slot) {
  // each lane owns CPL adjacent columns (8 or 16 bytes per variable plane);
  // lanes 0 and 63 are halo, so a wave produces 62*CPL columns
  constexpr int COLS = 62 * CPL;

  // ---- step control (dt from the previous step's reduction) -----------------
  double dt = 0.0;
  const int p = a.parity, q = p ^ 1;
  if (MODE != MODE_RESIDUAL) {
    const Ctrl* c = a.ctrl;
    const double t = c->t[p];
    const bool active = t < a.t_end;
    unsigned long long m = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) m = c->cmax[p][k].v > m ? c->cmax[p][k].v : m;
    // dt = 0.5*dx/cmax*cfl (benchmark_2d.f90:242)
    dt = 0.5 * a.h / __longlong_as_double((long long)m) * a.cfl;
    if (tile == 0 && lane < 8) {
      Ctrl* cw = a.ctrl;
      // a step past t_end changes nothing and carries the wave-speed maximum to
      // the next parity, so a later call with a larger t_end resumes from it
      if (a.first) cw->cmax[q][lane].v = active ? 0ull : c->cmax[p][lane].v;
      if (LAST && lane == 0) {
        cw->t[q] = active ? t + dt : t;
        if (active) { cw->steps_done += 1; cw->dt_last = dt; }
      }
    }
    if (!active) return;
  }

  const int tx = tile % a.ntx, ty = tile / a.ntx;
  const int nx = a.nx, ny = a.ny;
  const int xw = tx * COLS - CPL;              // first column of lane 0 (uniform)
  const int x0 = xw + CPL * lane;              // first column of this lane's block
  const int y0 = ty * a.rows;
  const int y1 = min(y0 + a.rows, ny);
  const bool own = lane >= 1 && lane <= 62 && x0 < nx;
  // interior strips read whole blocks; strips touching a side wall build their
  // halo column by column (mirror for walls, clamp for the frozen ring)
  const bool fast = xw >= 0 && xw + 64 * CPL <= nx;
  const long long plane = a.plane;
  const int pitch = a.pitch;
  // one variable's row holds min(pitch, plane) elements: pitch when the
  // variables are separate planes (plane = pitch * ny), plane when their rows
  // lie side by side (pitch = 4 * plane); every row access is bounded to it
  const int rowl = (long long)pitch < plane ? pitch : (int)plane;
  const int rowb = rowl * (int)sizeof(T);
  const unsigned vboff = (unsigned)clampi(x0, 0, rowl - CPL) * (unsigned)sizeof(T);

  auto load_row = [&](const T* base, int j) -> Blk<T, CPL> {
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
    Blk<T, CPL> b;
    if (fast) {
      Lanes<T, CPL> v0 = ldv<T, CPL>(r, vboff, rowb), v1 = ldv<T, CPL>(r + plane, vboff, rowb),
                    v2 = ldv<T, CPL>(r + 2 * plane, vboff, rowb), v3 = ldv<T, CPL>(r + 3 * plane, vboff, rowb);
#pragma unroll
      for (int k = 0; k < CPL; ++k) b.c[k] = Vec4<T>{v0.v[k], v1.v[k], v2.v[k], v3.v[k]};
    } else {
#pragma unroll
      for (int k = 0; k < CPL; ++k) {
        int col = x0 + k, mc;
        T sxk = T(1);
        if (BC == BC_REFLECT) {
          mc = col < 0 ? -1 - col : (col >= nx ? 2 * nx - 1 - col : col);
          if (col < 0 || col >= nx) sxk = T(-1);
          mc = clampi(mc, 0, nx - 1);
        } else {
          mc = clampi(col, 0, nx - 1);
        }
        const unsigned o = (unsigned)mc * (unsigned)sizeof(T);
        b.c[k] = Vec4<T>{ld1(r, o, rowb), sxk * ld1(r + plane, o, rowb), ld1(r + 2 * plane, o, rowb),
                         ld1(r + 3 * plane, o, rowb)};
      }
    }
    if (BC == BC_REFLECT && flip) {
#pragma unroll
      for (int k = 0; k < CPL; ++k) b.c[k].my = -b.c[k].my;
    }
    return b;
  };
  // interior strips: every row and column read lies inside the grid, so a row
  // is four plain block loads with nothing applied to the loaded values (no
  // mirror, no per-column path), which lets the loads stay in flight until the
  // row that uses them
  auto load_row_interior = [&](const T* base, int j) -> Blk<T, CPL> {
    const T* r = base + (long long)clampi(j, 0, ny - 1) * pitch;
    Lanes<T, CPL> v0 = ldv<T, CPL>(r, vboff, rowb), v1 = ldv<T, CPL>(r + plane, vboff, rowb),
                  v2 = ldv<T, CPL>(r + 2 * plane, vboff, rowb), v3 = ldv<T, CPL>(r + 3 * plane, vboff, rowb);
    Blk<T, CPL> b;
#pragma unroll
    for (int k = 0; k < CPL; ++k) b.c[k] = Vec4<T>{v0.v[k], v1.v[k], v2.v[k], v3.v[k]};
    return b;
  };
  // U^n rows: only owned cells of the tile's own rows use them, so halo lanes,
  // lanes past nx and the ring's row beyond the tile read 0 through an offset
  // past the extent instead of memory (0.7 % faster in FP32, no result changes)
  auto load_cell = [&](const T* base, int j) -> Blk<T, CPL> {
    const T* r = base + (long long)clampi(j, 0, ny - 1) * pitch;
    const unsigned uo = (own && j >= y0 && j < y1) ? vboff : 0x80000000u;
    Lanes<T, CPL> v0 = ldv<T, CPL>(r, uo, rowb), v1 = ldv<T, CPL>(r + plane, uo, rowb),
                  v2 = ldv<T, CPL>(r + 2 * plane, uo, rowb), v3 = ldv<T, CPL>(r + 3 * plane, uo, rowb);
    Blk<T, CPL> b;
#pragma unroll
    for (int k = 0; k < CPL; ++k) b.c[k] = Vec4<T>{v0.v[k], v1.v[k], v2.v[k], v3.v[k]};
    return b;
  };

  // a face state D + h*s of a reconstruction (h = +-1/2).  The deviation form
  // writes it as one fused multiply-add per variable, so every launch shape
  // rounds it the same way: h*s is exact for normal numbers, but not for
  // subnormal ones, where a separate product would round first.  The FP64
  // form keeps the reference's product-then-sum.
  auto face_state = [](const Vec4<T>& D, T h, const Vec4<T>& s) -> Vec4<T> {
    if constexpr (FORM == FORM_DEV) {
      auto fm = [](T x, T y, T z) -> T {
        if constexpr (sizeof(T) == 4) return __builtin_fmaf(x, y, z);
        else return __builtin_fma(x, y, z);
      };
      return {fm(h, s.r, D.r), fm(h, s.mx, D.mx), fm(h, s.my, D.my), fm(h, s.e, D.e)};
    } else {
      return D + scale(h, s);
    }
  };

  Phys<T> P;
  P.gamma = T(a.gamma);
  P.gm1 = T(a.gamma - 1.0);
  P.half = T(0.5);
  P.g = T(a.g);
  P.ggm1 = T(a.gamma * (a.gamma - 1.0));
  P.gfloor = T(a.gamma * 1e-10);
  const T odx = T(a.odx), ody = T(a.ody);
  // stage combination out = alpha*U^n + beta*D + (beta*dt)*L
  const T alpha = T(a.alpha), beta = T(a.beta), bdt = T(a.beta * dt);
  // the deviation form folds beta*dt into the residual's factors
  const T kx = T(a.beta * dt * a.odx), ky = T(a.beta * dt * a.ody), kg = T(a.beta * dt * a.g);

  // separable FP64 equilibrium: rho = (rho0 e^{-kx}) e^{-ky} is formed in FP64
  // and rounded once; p, E and E+p scale with rho by constant ratios.
  double rxf[CPL], rxc[CPL];
#pragma unroll
  for (int k = 0; k < CPL; ++k) {
    rxf[k] = a.rxf[clampi(x0 + k, 0, nx)];        // left x face of column x0+k
    rxc[k] = a.rxc[clampi(x0 + k, 0, nx - 1)];
  }
  const T por = T(a.p_over_rho), eor = T(a.p_over_rho * a.inv_gm1);
  const T epor = T(a.p_over_rho * a.gamma * a.inv_gm1);
  const T gpor = T(a.p_over_rho * a.gamma);
  auto face_eq = [&](double rx, double ey) -> FaceEq<T> {
    FaceEq<T> f;
    f.rho = T(rx * ey);
    f.p = f.rho * por;
    f.e = f.rho * eor;
    f.ep = f.rho * epor;
    f.gp = f.rho * gpor;
    return f;
  };

  // the wave-speed maximum in the working precision (exact when widened)
  T smax = T(0);
  bool bad = false;

  // ---- the sweep ------------------------------------------------------------------
  // Even tile rows sweep upward, odd ones downward, so two vertically adjacent
  // tiles reach their shared halo rows at the same time and one of them finds
  // the rows in L2.  Both directions compute every face flux from the same two
  // states in the same order (lower, upper), so the result does not depend on
  // the direction.
  // A ring of RING rows holds rows j, j+DIR, j+2DIR while row j is computed and
  // receives row j+(RING-1)DIR (one row in flight ahead of use); U^n uses a ring
  // of URING.  The body is unrolled RING times with rotated roles, so the rings
  // never move data.  (A deeper ring measured no faster: the strip sweep is
  // limited by its memory access pattern, not by loads in flight per wave.)
  auto sweep = [&](auto dirc, auto interiorc) {
    constexpr int DIR = decltype(dirc)::value;
    auto load_row_s = [&](const T* base, int j) -> Blk<T, CPL> {
      if constexpr (decltype(interiorc)::value) return load_row_interior(base, j);
      else return load_row(base, j);
    };
    constexpr int RING = 4;
    // FP64 holds U^n for one row only (loaded as the row starts, used as it
    // ends): that keeps its stage-2/3 kernel under 256 VGPRs, two waves per
    // SIMD instead of one.  FP32 keeps two rows at three waves, which measured
    // faster than one row at four.
    constexpr int URING = sizeof(T) == 8 ? 1 : 2;
    // slopes are differences along the sweep (minmod2(-a,-b) = -minmod2(a,b)), so a
    // row's state toward the next row is D + sy/2 and toward the previous D - sy/2
    const T hd = P.half;
    const int js = DIR > 0 ? y0 : y1 - 1;      // first row of the sweep
    Blk<T, CPL> R[RING], Ur[URING], sy, Gp;
    // flux through the face between row jb (behind, state Db) and row ja (ahead, Da)
    auto yface = [&](const FaceEq<T>& qy, const Vec4<T>& Db, const Vec4<T>& Da) {
      if constexpr (DIR > 0) return face_flux<T, FLUX, 1, FORM>(P, qy, Db, Da);
      else return face_flux<T, FLUX, 1, FORM>(P, qy, Da, Db);
    };
    {
      const double eyf0 = table_at(a.eyf, clampi(js + (DIR < 0), 0, ny));
      Blk<T, CPL> Dm1 = load_row_s(a.in, js - DIR);
#pragma unroll
      for (int k = 0; k < RING - 1; ++k) R[k] = load_row_s(a.in, js + k * DIR);
      Blk<T, CPL> Dm2;
      if (RECON) Dm2 = load_row_s(a.in, js - 2 * DIR);
#pragma unroll
      for (int k = 0; k < CPL; ++k) {
        FaceEq<T> qy = face_eq(rxc[k], eyf0);
        if (RECON) {
          Vec4<T> sym1 = minmod2(Dm1.c[k] - Dm2.c[k], R[0].c[k] - Dm1.c[k]);
          sy.c[k] = minmod2(R[0].c[k] - Dm1.c[k], R[1].c[k] - R[0].c[k]);
          Gp.c[k] = yface(qy, face_state(Dm1.c[k], hd, sym1), face_state(R[0].c[k], -hd, sy.c[k]));
        } else {
          Gp.c[k] = yface(qy, Dm1.c[k], R[0].c[k]);
        }
      }
#pragma unroll
      for (int k = 0; k < URING; ++k) Ur[k] = Blk<T, CPL>{};
      if (MODE == MODE_STAGE) {
#pragma unroll
        for (int k = 0; k < URING - 1; ++k) Ur[k] = load_cell(a.un, js + k * DIR);
      }
    }

    // one row: A = row j, B = row j+DIR, C = row j+2DIR; S receives row
    // j+(RING-1)DIR; UA holds U^n(j) and US receives U^n(j+(URING-1)DIR).
    auto row_step = [&](int j, const Blk<T, CPL>& A, const Blk<T, CPL>& B, const Blk<T, CPL>& C,
                        Blk<T, CPL>& S, const Blk<T, CPL>& UA, Blk<T, CPL>& US) {
      S = load_row_s(a.in, j + (RING - 1) * DIR);
      if (MODE == MODE_STAGE) US = load_cell(a.un, j + (URING - 1) * DIR);
      const int jt = clampi(j, 0, ny - 1);
      const double eyc = table_at(a.eyc, jt);
      const double eyf = table_at(a.eyf, jt + (DIR > 0));

      // y faces between row j and row j+DIR
      Blk<T, CPL> Gn;
#pragma unroll
      for (int k = 0; k < CPL; ++k) {
        FaceEq<T> qy = face_eq(rxc[k], eyf);
        if (RECON) {
          Vec4<T> syn = minmod2(B.c[k] - A.c[k], C.c[k] - B.c[k]);
          Gn.c[k] = yface(qy, face_state(A.c[k], hd, sy.c[k]), face_state(B.c[k], -hd, syn));
          sy.c[k] = syn;
        } else {
          Gn.c[k] = yface(qy, A.c[k], B.c[k]);
        }
      }
      // x faces on the left of each column; the first takes its left state from
      // the neighbouring lane's last column, the last column's right face is the
      // right neighbour's first
      Vec4<T> xr[CPL], xl[CPL];
      if (RECON) {
        Vec4<T> lm = from_left(A.c[CPL - 1]), rp = from_right(A.c[0]);
#pragma unroll
        for (int k = 0; k < CPL; ++k) {
          Vec4<T> am = k == 0 ? lm : A.c[k - 1];
          Vec4<T> ap = k == CPL - 1 ? rp : A.c[k + 1];
          Vec4<T> sx = minmod2(A.c[k] - am, ap - A.c[k]);
          xr[k] = face_state(A.c[k], P.half, sx);
          xl[k] = face_state(A.c[k], -P.half, sx);
        }
      } else {
#pragma unroll
        for (int k = 0; k < CPL; ++k) xr[k] = xl[k] = A.c[k];
      }
      Vec4<T> F[CPL + 1];
      const Vec4<T> xr_left = from_left(xr[CPL - 1]);
#pragma unroll
      for (int k = 0; k < CPL; ++k)
        F[k] = face_flux<T, FLUX, 0, FORM>(P, face_eq(rxf[k], eyc), k == 0 ? xr_left : xr[k - 1], xl[k]);
      F[CPL] = from_right(F[0]);

      Blk<T, CPL> O;
#pragma unroll
      for (int k = 0; k < CPL; ++k) {
        const Vec4<T>& D = A.c[k];
        // residual: -dF/dx - dG/dy + (s(U) - s(U_eq)), s = (0, -rho g, -rho g, -g (m_x+m_y));
        // the same expression in both directions, on the flux above (Ga) and below (Gb)
        const Vec4<T>& Ga = DIR > 0 ? Gn.c[k] : Gp.c[k];
        const Vec4<T>& Gb = DIR > 0 ? Gp.c[k] : Gn.c[k];
        const bool frozen =
            BC == BC_FROZEN && (x0 + k == 0 || x0 + k == nx - 1 || j == 0 || j == ny - 1);
        if (FORM == FORM_DEV && MODE != MODE_RESIDUAL) {
          // out = base - kx dF - ky dG + kg s, base = D (stage 1) or alpha U^n + beta D
          Vec4<T> B = D;
          if (MODE == MODE_STAGE) B = scale(alpha, UA.c[k]) + scale(beta, D);
          Vec4<T> Q;
          Q.r = B.r - kx * (F[k + 1].r - F[k].r) - ky * (Ga.r - Gb.r);
          Q.mx = B.mx - kx * (F[k + 1].mx - F[k].mx) - ky * (Ga.mx - Gb.mx) - kg * D.r;
          Q.my = B.my - kx * (F[k + 1].my - F[k].my) - ky * (Ga.my - Gb.my) - kg * D.r;
          Q.e = B.e - kx * (F[k + 1].e - F[k].e) - ky * (Ga.e - Gb.e) - kg * (D.mx + D.my);
          O.c[k] = frozen ? B : Q;
        } else {
          Vec4<T> L;
          L.r = -(F[k + 1].r - F[k].r) * odx - (Ga.r - Gb.r) * ody;
          L.mx = (-(F[k + 1].mx - F[k].mx) * odx - (Ga.mx - Gb.mx) * ody) - P.g * D.r;
          L.my = (-(F[k + 1].my - F[k].my) * odx - (Ga.my - Gb.my) * ody) - P.g * D.r;
          L.e = (-(F[k + 1].e - F[k].e) * odx - (Ga.e - Gb.e) * ody) - P.g * (D.mx + D.my);
          if (frozen) L = {T(0), T(0), T(0), T(0)};
          if (MODE == MODE_RESIDUAL) {
            O.c[k] = L;
          } else if (MODE == MODE_STAGE1) {
            O.c[k] = D + scale(bdt, L);
          } else {
            O.c[k] = scale(alpha, UA.c[k]) + (scale(beta, D) + scale(bdt, L));
          }
        }
      }
      Gp = Gn;
      // The stores are issued on every row and lane, without a branch: a lane
      // or row that owns nothing gets an offset past the row's extent, which
      // the buffer range check discards.  A branch around them would make the
      // wait for the next row's loads (issued before these stores) also wait
      // for the stores, since vector loads and stores share one in-order
      // counter and the two paths would be merged pessimistically.
      const bool row_ok = j >= y0 && j < y1;
      {
        T* d = a.out + (long long)jt * pitch;
        const unsigned soff = (own && row_ok) ? vboff : 0x80000000u;
        Lanes<T, CPL> o0, o1, o2, o3;
#pragma unroll
        for (int k = 0; k < CPL; ++k) {
          o0.v[k] = O.c[k].r; o1.v[k] = O.c[k].mx; o2.v[k] = O.c[k].my; o3.v[k] = O.c[k].e;
        }
        // a block may run past nx into the row padding, which is never read as data
        stv<T, CPL>(d, soff, rowb, o0);
        stv<T, CPL>(d + plane, soff, rowb, o1);
        stv<T, CPL>(d + 2 * plane, soff, rowb, o2);
        stv<T, CPL>(d + 3 * plane, soff, rowb, o3);
      }
      if (MODE != MODE_RESIDUAL && LAST && row_ok) {
#pragma unroll
        for (int k = 0; k < CPL; ++k) {
          if (own && x0 + k < nx) {
            bool ok;
            T rct = T(rxc[k] * eyc);
            T sp = cell_speed(P, rct, rct * eor, O.c[k], ok);
            smax = tmax(smax, sp);
            bad |= !ok;
          }
        }
      }
    };

    const int nrows = y1 - y0;
    for (int n = 0; n < nrows; n += RING) {
#pragma unroll
      for (int k = 0; k < RING; ++k)
        row_step(js + (n + k) * DIR, R[k % RING], R[(k + 1) % RING], R[(k + 2) % RING], R[(k + RING - 1) % RING],
                 Ur[k % URING], Ur[(k + URING - 1) % URING]);
    }
  };
  // rows read: y0-3 .. y1+2 (the ring runs one row past each end of the strip)
  const bool interior = fast && y0 >= 3 && y1 + 3 <= ny;
  if (ty & 1) {
    if (interior) sweep(std::integral_constant<int, -1>{}, std::true_type{});
    else sweep(std::integral_constant<int, -1>{}, std::false_type{});
  } else {
    if (interior) sweep(std::integral_constant<int, 1>{}, std::true_type{});
    else sweep(std::integral_constant<int, 1>{}, std::false_type{});
  }

  if (MODE != MODE_RESIDUAL && LAST) {
    if (bad) atomicOr(&a.ctrl->err, 1u);
    double sm = (double)smax;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) sm = fmax(sm, __shfl_xor(sm, off));
    if (lane == 0) atomicMax(&a.ctrl->cmax[q][slot].v, (unsigned long long)__double_as_longlong(sm));
  }
}

template <typename T, int RECON, int FLUX, int BC, int FORM, int MODE, int CPL, bool LAST>
__global__ __launch_bounds__(256) void fv2d_stage_kernel(StageArgs<T> a) {
  const int lane = threadIdx.x & 63;
  // wave index made explicitly uniform so that strip, row and row pointers
  // live in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tile = xcd_block(blockIdx.x, gridDim.x) * 4 + wave;
  if (tile >= a.ntiles) return;
  stage_tile<T, RECON, FLUX, BC, FORM, MODE, CPL, LAST>(a, tile, lane, blockIdx.x & 7);
}

// Persistent batched stage: every wave walks the global tile list of all
// instances (static stride, each XCD taking contiguous runs) and finds each
// tile's instance by binary search over the tile prefix sums.
template <typename T, int RECON, int FLUX, int BC, int FORM, int MODE, int CPL, bool LAST>
__global__ __launch_bounds__(256) void fv2d_batch_stage_kernel(BatchArgs<T> b) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int stride = gridDim.x * 4;
  for (int t = xcd_block(blockIdx.x, gridDim.x) * 4 + wave; t < b.ntiles; t += stride) {
    int lo = 0, hi = b.ninst - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (b.tile_start[mid] <= t) lo = mid; else hi = mid - 1;
    }
    const InstDesc& d = b.inst[lo];
    StageArgs<T> a;
    a.in = b.in + d.off;
    a.un = b.un ? b.un + d.off : nullptr;
    a.out = b.out + d.off;
    a.rxf = b.tabx + d.tx_off;
    a.rxc = b.tabx + d.tx_off + d.nx + 1;
    a.eyc = b.taby + d.ty_off;
    a.eyf = b.taby + d.ty_off + d.ny;
    a.ctrl = b.ctrl + lo;
    a.plane = d.plane;
    a.pitch = d.pitch; a.nx = d.nx; a.ny = d.ny;
    a.rows = d.rows; a.ntx = d.ntx; a.ntiles = d.ntiles;
    a.parity = b.parity; a.first = b.first; a.last = b.last;
    a.t_end = b.t_end; a.h = d.h; a.cfl = d.cfl;
    a.gamma = d.gamma; a.g = d.g; a.p_over_rho = d.p_over_rho; a.inv_gm1 = d.inv_gm1;
    a.odx = d.odx; a.ody = d.ody;
    a.alpha = b.alpha; a.beta = b.beta;
    stage_tile<T, RECON, FLUX, BC, FORM, MODE, CPL, LAST>(a, t - b.tile_start[lo], lane, blockIdx.x & 7);
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
  P.ggm1 = T(gamma * (gamma - 1.0)); P.gfloor = T(gamma * 1e-10);
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
    atomicMax(&ctrl->cmax[0][blockIdx.x & 7].v, (unsigned long long)__double_as_longlong(m));
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

template <typename T, int RECON, int FLUX, int BC, int FORM, int MODE, int CPL>
static hipError_t launch_stage_t(const StageArgs<T>& a, hipStream_t s) {
  int blocks = (a.ntiles + 3) / 4;
  if (MODE == MODE_STAGE && a.last)
    hipLaunchKernelGGL((fv2d_stage_kernel<T, RECON, FLUX, BC, FORM, MODE, CPL, true>), dim3(blocks), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((fv2d_stage_kernel<T, RECON, FLUX, BC, FORM, MODE, CPL, false>), dim3(blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

template <typename T, int FORM, int MODE, int CPL>
static hipError_t dispatch_scheme(const StageArgs<T>& a, int recon, int flux, int bc, hipStream_t s) {
#define FVWB_CASE(R, F, B)                                              \
  if (recon == R && flux == F && bc == B) return launch_stage_t<T, R, F, B, FORM, MODE, CPL>(a, s);
  FVWB_CASE(0, 0, 0) FVWB_CASE(0, 0, 1) FVWB_CASE(0, 1, 0) FVWB_CASE(0, 1, 1)
  FVWB_CASE(1, 0, 0) FVWB_CASE(1, 0, 1) FVWB_CASE(1, 1, 0) FVWB_CASE(1, 1, 1)
#undef FVWB_CASE
  return hipErrorInvalidValue;
}

template <typename T, int RECON, int FLUX, int BC, int FORM, int MODE, int CPL>
static hipError_t launch_batch_t(const BatchArgs<T>& b, int blocks, hipStream_t s) {
  if (MODE == MODE_STAGE && b.last)
    hipLaunchKernelGGL((fv2d_batch_stage_kernel<T, RECON, FLUX, BC, FORM, MODE, CPL, true>), dim3(blocks), dim3(256), 0,
                       s, b);
  else
    hipLaunchKernelGGL((fv2d_batch_stage_kernel<T, RECON, FLUX, BC, FORM, MODE, CPL, false>), dim3(blocks), dim3(256), 0,
                       s, b);
  return hipGetLastError();
}

template <typename T, int FORM, int MODE, int CPL>
static hipError_t dispatch_batch(const BatchArgs<T>& b, int blocks, int recon, int flux, int bc, hipStream_t s) {
#define FVWB_CASE(R, F, B)                                              \
  if (recon == R && flux == F && bc == B) return launch_batch_t<T, R, F, B, FORM, MODE, CPL>(b, blocks, s);
  FVWB_CASE(0, 0, 0) FVWB_CASE(0, 0, 1) FVWB_CASE(0, 1, 0) FVWB_CASE(0, 1, 1)
  FVWB_CASE(1, 0, 0) FVWB_CASE(1, 0, 1) FVWB_CASE(1, 1, 0) FVWB_CASE(1, 1, 1)
#undef FVWB_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_fv2d_batch_f64(const BatchArgs<double>& b, int blocks, int recon, int flux, int bc, int mode,
                                 hipStream_t s) {
  if (mode == MODE_STAGE1) return dispatch_batch<double, FORM_PLAIN, MODE_STAGE1, 2>(b, blocks, recon, flux, bc, s);
  if (mode == MODE_STAGE) return dispatch_batch<double, FORM_PLAIN, MODE_STAGE, 2>(b, blocks, recon, flux, bc, s);
  return hipErrorInvalidValue;
}

hipError_t launch_fv2d_batch_f32(const BatchArgs<float>& b, int blocks, int recon, int flux, int bc, int mode,
                                 int cpl, hipStream_t s) {
  if (cpl == 4) {
    if (mode == MODE_STAGE1) return dispatch_batch<float, FORM_DEV, MODE_STAGE1, 4>(b, blocks, recon, flux, bc, s);
    if (mode == MODE_STAGE) return dispatch_batch<float, FORM_DEV, MODE_STAGE, 4>(b, blocks, recon, flux, bc, s);
  } else if (cpl == 2) {
    if (mode == MODE_STAGE1) return dispatch_batch<float, FORM_DEV, MODE_STAGE1, 2>(b, blocks, recon, flux, bc, s);
    if (mode == MODE_STAGE) return dispatch_batch<float, FORM_DEV, MODE_STAGE, 2>(b, blocks, recon, flux, bc, s);
  }
  return hipErrorInvalidValue;
}

// Stage modes: 0 residual (FP64 plain only), 1 first stage, 2 later stages.
// FP64 lanes own 2 columns (16-byte loads); FP32 lanes own `cpl` = 2 or 4.
hipError_t launch_fv2d_stage_f64(const StageArgs<double>& a, int recon, int flux, int bc, int mode,
                                 hipStream_t s) {
  if (mode == MODE_RESIDUAL) return dispatch_scheme<double, FORM_PLAIN, MODE_RESIDUAL, 2>(a, recon, flux, bc, s);
  if (mode == MODE_STAGE1) return dispatch_scheme<double, FORM_PLAIN, MODE_STAGE1, 2>(a, recon, flux, bc, s);
  return dispatch_scheme<double, FORM_PLAIN, MODE_STAGE, 2>(a, recon, flux, bc, s);
}

hipError_t launch_fv2d_stage_f32(const StageArgs<float>& a, int recon, int flux, int bc, int mode, int cpl,
                                 hipStream_t s) {
  if (cpl == 4) {
    if (mode == MODE_STAGE1) return dispatch_scheme<float, FORM_DEV, MODE_STAGE1, 4>(a, recon, flux, bc, s);
    if (mode == MODE_STAGE) return dispatch_scheme<float, FORM_DEV, MODE_STAGE, 4>(a, recon, flux, bc, s);
  } else if (cpl == 2) {
    if (mode == MODE_STAGE1) return dispatch_scheme<float, FORM_DEV, MODE_STAGE1, 2>(a, recon, flux, bc, s);
    if (mode == MODE_STAGE) return dispatch_scheme<float, FORM_DEV, MODE_STAGE, 2>(a, recon, flux, bc, s);
  }
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

// Plain streaming copy (16 B per lane): the calibration point for
// the achieved-bandwidth and HBM-counter figures of the solver kernels.
__global__ __launch_bounds__(256) void copy_probe_kernel(const float4* __restrict__ src, float4* __restrict__ dst,
                                                         long long n4) {
  // one float4 per thread over a grid that covers the buffer: the fastest plain
  // device copy pattern on MI355X (measured against grid-stride variants)
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n4) dst[k] = src[k];
}

hipError_t launch_copy_probe(const void* src, void* dst, long long bytes, hipStream_t s) {
  const long long n4 = bytes / 16;
  const long long blocks = (n4 + 255) / 256;
  if (blocks < 1 || blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(copy_probe_kernel, dim3((unsigned)blocks), dim3(256), 0, s, (const float4*)src, (float4*)dst, n4);
  return hipGetLastError();
}

}  // namespace fvwb
