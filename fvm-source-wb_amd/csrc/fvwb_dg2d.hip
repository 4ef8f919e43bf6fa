// fvwb_dg2d.hip — 2D modal DG (P2, 3x3 Gauss-Legendre) in deviation form on MI355X.
//
// Residual of 2d/benchmark_2d_dg.f90:1137-1479 (volume integrals :1207-1244,
// traces :1253-1314, face fluxes :1333-1366, edge integrals :1372-1411, source
// :1426-1443, assembly :1455-1462) written for the deviation dU = U - U_eq of
// dg_with_source.f90:1749-2031: all fluxes and the source are deviations from
// the equilibrium's, so a resting atmosphere is an exact fixed point.
//
// Work decomposition: one wave sweeps a strip of 16 columns (14 owned plus one
// halo column on each side) upward through a segment of rows, one row of 16
// cells per step.  The projections of a row are small dense products and run on
// the matrix cores as v_mfma_f32_16x16x4f32 with the 16 cells in the N
// dimension:
//   forward:  [16 points x 8 modes] x [8 modes x 16 cells], once for the 9
//             volume nodes and once for the 12 face points;
//   backward: [16 outputs x 4 inputs] x [4 inputs x 16 cells] per step, 8
//             steps per variable for the 18 node flux terms and the 12 face
//             flux terms; the backward operands carry dt/dx or dt/dy.
// Lane l = 16 g + c handles cell c; in the forward products it receives two
// volume nodes (three in group 3) and the left, right, bottom and top traces
// at Gauss point g (g = 0..2); in the backward products it feeds exactly those
// values' fluxes as k-slot g, so the point-wise flux work stays in the lane and
// all four k-slots carry data.  The output rows of the
// backward products are ordered so that lane (g, c) receives modes g and g+4 of
// cell c: the layout in which it loaded them, so the stage combination needs no
// data movement.  x neighbours are one DPP row shift away; the trace of the row
// below is carried in registers from the previous step.  No LDS, no barriers.
//
// The source deviation -g (0, drho, drho, dmx + dmy) is linear in the state and
// the 3-point Gauss rule integrates phi_i phi_j exactly, so its projection is
// -g times the modes themselves and is added directly.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/fvwb.h"
#include "fvwb_checkpoint.h"
#include "fvwb_device.h"

namespace fvwb {
namespace dg {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int NM = 6;      // P2 modes
constexpr int NQ = 3;      // Gauss points per direction
constexpr int SW = 14;     // owned columns per strip (16 lanes, one halo column each side)
constexpr int GROWS = 2;   // ghost rows (one below row 0, one above row ny-1)
constexpr unsigned MASKED = 0x80000000u;   // buffer offset past every extent: loads return 0, stores drop

// Per-lane MFMA operand constants, built on the host from the basis.  Output row
// 4g + s of a backward product is mode g (s = 0) or mode g + 4 (s = 1, g < 2).
//
// The nine volume nodes are spread over all four lane groups: group g takes
// node n = 2g + t in step t = 0, 1 (g < 3) and group 3 takes n = 6 + t in
// steps t = 0, 1, 2, where node n is (x_{n%3}, y_{n/3}).  Every group then
// feeds at most seven values per variable into the backward products of its
// row (groups 0-2: F_x, F_y of two nodes and the left, right and bottom face
// fluxes; group 3: F_x, F_y of three nodes), so seven products per variable do,
// plus the top face's with the row's finish.
struct Consts {
  float fwd_node[16][8];        // [row 4g + t: node of group g in step t][mode]; rows without a node are zero
  float fwd_face[16][8];        // [row 4q + s: point q, s = 0 left, 1 right, 2 bottom, 3 top][mode]
  float back[7][16][4];         // [slot][output row][k = lane group of the value]
  float back_top[16][4];        // top face at point k (k < 3)
};

// the volume node of lane group g in step t (see Consts)
__host__ __device__ constexpr int vol_node(int g, int t) { return g < 3 ? 2 * g + (t < 1 ? t : 1) : 6 + t; }

// The CFL rate max((|u|+c)/dx + (|v|+c)/dy) of parity p is the maximum of
// rate[p][0..RSLOTS-1]: workgroup b reduces into slot b % RSLOTS, each slot on a
// 128-byte line of its own, so the grid's atomics spread over RSLOTS lines
// instead of queueing on one.
constexpr int RSLOTS = 16;
struct alignas(128) RateSlot {
  uint32_t v;                 // FP32 bit pattern; positive floats order like uints
  uint32_t pad[31];
};

struct Ctrl {
  RateSlot rate[2][RSLOTS];
  double t[2];
  double dt_last;
  int32_t steps;
  uint32_t err;       // bit 0: rho<=0, p<=0 or non-finite seen
};

struct Args {
  const float* in;     // stage input planes (24 = 4 vars x 6 modes), base of the ghost frame
  const float* un;     // U^n planes (stage combination), may be null when alpha == 0
  float* out;          // output planes (may alias un)
  const Consts* C;
  const double* exn;   // e^{-k x} at x nodes   [3*nx]  (FP64)
  const double* eyn;   // e^{-k y} at y nodes   [3*ny]
  const double* exf;   // e^{-k x} at x faces   [nx+1]
  const double* eyf;   // e^{-k y} at y faces   [ny+1]
  Ctrl* ctrl;
  long long plane;
  int pitch, nx, ny, parity;
  int ntx, nseg, seg_rows;   // strips across x, segments along y, rows per segment
  int first, residual_only;
  double t_end, cfl_over;    // cfl / (2k+1)
  double dtc;                // this stage's multiple of the step's dt
  float alpha, beta;
  // the two-output stage (DUAL): out = alpha U^n + beta U + gam dt L(U) and
  // xout = pu U^n + px X + pin U + pl dt L(U), X read from xin (own cells only)
  const float* xin;
  float* xout;
  float gam, pu, px, pin, pl;
  float gamma, g, odx, ody;
  double rho0;               // equilibrium density per e^{-k(x+y)}
  float por, eor, epor, gpor;   // p, E = p/(gamma-1), E+p and gamma p of the equilibrium per unit density
  float ggm1, gfloor;        // gamma (gamma-1) and gamma*1e-10
};

// A table entry at a wave-uniform index, read through the constant address
// space so that it becomes a scalar load.
__device__ __forceinline__ double ctab(const double* t, int k) {
  return ((const __attribute__((address_space(4))) double*)t)[k];
}

__device__ inline float ctrl_rate(const Ctrl* c, int par) {
  uint32_t m = 0u;
#pragma unroll
  for (int k = 0; k < RSLOTS; ++k) m = max(m, c->rate[par][k].v);
  return __uint_as_float(m);
}

// Deviation fluxes along x and y at one point.
__device__ inline void point_terms(const Phys<float>& P, const FaceEq<float>& q, Vec4<float> d, Vec4<float>& fx,
                                   Vec4<float>& fy) {
  float rho = q.rho + d.r;
  float inv = trcp(tmax(rho, 1e-10f));
  float u = d.mx * inv, v = d.my * inv;
  float dp = P.gm1 * (d.e - P.half * (d.mx * u + d.my * v));
  float ep = (q.ep + d.e) + dp;
  fx = {d.mx, d.mx * u + dp, d.my * u, ep * u};
  fy = {d.my, d.mx * v, d.my * v + dp, ep * v};
}

// Equilibrium at a point with e = e^{-kx} e^{-ky} from the separable FP64
// tables: the density is formed in FP64 and rounded once; p, E (HLLC only),
// E+p and gamma p are constant multiples of it (the isothermal atmosphere).
__device__ inline FaceEq<float> eq_at(const Args& a, double e) {
  FaceEq<float> q;
  q.rho = (float)(a.rho0 * e);
  q.p = q.rho * a.por;
  q.e = q.rho * a.eor;
  q.ep = q.rho * a.epor;
  q.gp = q.rho * a.gpor;
  return q;
}

// DPP moves inside each row of 16 lanes (one row = the 16 cells of a strip):
// lane c receives lane c-1 (c+1); the lane at the row's edge receives 0.
__device__ __forceinline__ float from_col_left(float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x111 /*row_shr:1*/,
                                                               0xF, 0xF, true));
}
__device__ __forceinline__ float from_col_right(float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x101 /*row_shl:1*/,
                                                               0xF, 0xF, true));
}

// Buffer resource over one row of one variable's six mode planes; offsets past
// nrec bytes read 0 and drop stores.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t mode_rsrc(const float* p, int nrec) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)p, 0, nrec, 0x00020000);
}
__device__ __forceinline__ float ldm(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ void stm(__amdgpu_buffer_rsrc_t r, unsigned off, float x) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, off, 0, 0);
}

__device__ __forceinline__ v4f mfma(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Blocks are dealt round-robin to the 8 XCDs; renumber them so that each XCD
// receives a contiguous range of strips (neighbouring halo columns share its L2).
__device__ __forceinline__ int xcd_block(int b, int nb) {
  int xcd = b & 7, idx = b >> 3, base = nb >> 3, rem = nb & 7;
  return xcd * base + (xcd < rem ? xcd : rem) + idx;
}

template <int FLUX, bool DUAL>
__global__ __launch_bounds__(256) void dg_stage_kernel(Args a) {
  const Ctrl* ck = a.ctrl;
  const int par = a.parity;
  double dt = 0.0;
  if (!a.residual_only) {
    const double t = ck->t[par];
    if (!(t < a.t_end)) return;       // this instance is done
    dt = a.cfl_over / (double)ctrl_rate(ck, par);
    if (t + dt > a.t_end) dt = a.t_end - t;
    if (a.first && blockIdx.x == 0 && threadIdx.x < RSLOTS) a.ctrl->rate[par ^ 1][threadIdx.x].v = 0u;
  }
  const int w = xcd_block(blockIdx.x, gridDim.x) * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (w >= a.ntx * a.nseg) return;
  const int strip = w % a.ntx, seg = w / a.ntx;
  const int y0 = seg * a.seg_rows, y1 = min(y0 + a.seg_rows, a.ny);
  const int lane = threadIdx.x & 63, c = lane & 15, g = lane >> 4;
  const int gq = min(g, 2);                   // Gauss point of this lane group (group 3 idles)
  const int i = strip * SW - 1 + c;           // column of this lane
  const bool own = c >= 1 && c <= SW && i < a.nx;
  const int nx = a.nx, ny = a.ny;
  const long long plane = a.plane;
  const int pitch = a.pitch;
  // accumulator acc = U + dtc dt L (the backward operands carry the factor
  // dtc dt), out = alpha U^n + beta acc; the residual alone is acc with dt = 1
  // and no U; the two-output stage accumulates dt L alone
  const float sdt = a.residual_only ? 1.f : (float)(DUAL ? dt : a.dtc * dt);
  const float ku = (a.residual_only || DUAL) ? 0.f : 1.f;
  const float gdt = a.g * sdt;
  const float alpha = a.residual_only ? 0.f : a.alpha;
  const float beta = a.residual_only ? 1.f : a.beta;
  const bool with_un = !a.residual_only && a.alpha != 0.f;

  Phys<float> P;
  P.gamma = a.gamma; P.gm1 = a.gamma - 1.0f; P.half = 0.5f; P.g = a.g;
  P.ggm1 = a.ggm1; P.gfloor = a.gfloor;

  // lane offsets (bytes) of modes g and g+4 of column i inside a row of a variable
  const unsigned colb = (unsigned)(i + 1) * 4u;
  const unsigned off0 = (unsigned)g * (unsigned)plane * 4u + colb;
  const unsigned off1 = g + 4 < NM ? (unsigned)(g + 4) * (unsigned)plane * 4u + colb : MASKED;
  const unsigned st0 = own ? off0 : MASKED, st1 = own ? off1 : MASKED;
  const int nrec = (int)((5 * plane + pitch) * 4);
  auto rs = [&](const float* base, int v, int j) {
    return mode_rsrc(base + (long long)(v * NM) * plane + (long long)(j + 1) * pitch, nrec);
  };
  auto load_row = [&](const float* base, int j, float (&B)[4][2]) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const __amdgpu_buffer_rsrc_t r = rs(base, v, j);
      B[v][0] = ldm(r, off0);
      B[v][1] = ldm(r, off1);
    }
  };

  // U^n of the owned cells only (halo lanes read 0)
  auto load_own = [&](const float* base, int j, float (&B)[4][2]) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const __amdgpu_buffer_rsrc_t r = rs(base, v, j);
      B[v][0] = ldm(r, st0);
      B[v][1] = ldm(r, st1);
    }
  };
  // the same, reading 0 everywhere when ok is false (base must still be a
  // buffer of the frame: without U^n it is the input)
  const float* un_base = with_un ? a.un : a.in;
  auto load_own_if = [&](bool ok, const float* base, int j, float (&B)[4][2]) {
    const unsigned o0 = ok ? st0 : MASKED, o1 = ok ? st1 : MASKED;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const __amdgpu_buffer_rsrc_t r = rs(base, v, j);
      B[v][0] = ldm(r, o0);
      B[v][1] = ldm(r, o1);
    }
  };

  // MFMA operands
  const Consts& C = *a.C;
  const float An0 = C.fwd_node[c][g], An1 = C.fwd_node[c][g + 4];
  const float Af0 = C.fwd_face[c][g], Af1 = C.fwd_face[c][g + 4];
  // backward slots: 0-3 take F_x, F_y, F_x, F_y of steps 0 and 1; 4 and 5 the
  // left and right face fluxes (groups 0-2) or step 2's F_x and F_y (group 3);
  // 6 the bottom face flux (groups 0-2).  x-direction terms carry dt/dx.
  float Ab[7];
#pragma unroll
  for (int s = 0; s < 7; ++s) {
    const bool xd = s == 0 || s == 2 || s == 4 || (s == 5 && g < 3);
    Ab[s] = (sdt * (xd ? a.odx : a.ody)) * C.back[s][c][g];
  }
  const float Abt = (sdt * a.ody) * C.back_top[c][g];

  // equilibrium factors of this column: the x factor of this group's node in
  // each step, the left x face, x node gq; the y index of each step's node
  const int ic = min(max(i, 0), nx - 1), ifc = min(max(i, 0), nx);
  double exr[3];
  int nyb[3];
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    const int n = vol_node(g, t);
    exr[t] = a.exn[3 * ic + n % 3];
    nyb[t] = n / 3;
  }
  const double exf_i = a.exf[ifc];
  const double exg = a.exn[3 * ic + gq];

  auto fwd_face = [&](const float (&B)[4][2], v4f (&df)[4]) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const v4f z = {0.f, 0.f, 0.f, 0.f};
      df[v] = mfma(Af0, B[v][0], z);
      df[v] = mfma(Af1, B[v][1], df[v]);
    }
  };
  auto mirror_x = [](Vec4<float> d) { return Vec4<float>{d.r, -d.mx, d.my, d.e}; };
  auto mirror_y = [](Vec4<float> d) { return Vec4<float>{d.r, d.mx, -d.my, d.e}; };
  // y-face flux at Gauss point gq below row j (lower state dl, upper dr)
  auto flux_y = [&](int j, Vec4<float> dl, Vec4<float> dr) {
    return face_flux<float, FLUX, 1, FORM_DEV>(P, eq_at(a, exg * a.eyf[j]), dl, dr);
  };
  // finish row j: add the flux through its top face, combine and store
  float Uo[4][2], Xo[4][2];   // the two-output stage: U and X of the row being finished
  // (ok false: the row below the segment's first row, whose stores are dropped
  // through their offsets rather than skipped by a branch)
  auto finish = [&](int j, bool ok, v4f (&acc)[4], const Vec4<float>& Ft, const float (&N)[4][2]) {
    const float ft[4] = {Ft.r, Ft.mx, Ft.my, Ft.e};
    const unsigned o0 = ok ? st0 : MASKED, o1 = ok ? st1 : MASKED;
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[v] = mfma(Abt, ft[v], acc[v]);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const __amdgpu_buffer_rsrc_t r = rs(a.out, v, j);
      if constexpr (!DUAL) {
        stm(r, o0, alpha * N[v][0] + beta * acc[v][0]);
        stm(r, o1, alpha * N[v][1] + beta * acc[v][1]);
      } else {
        stm(r, o0, alpha * N[v][0] + beta * Uo[v][0] + a.gam * acc[v][0]);
        stm(r, o1, alpha * N[v][1] + beta * Uo[v][1] + a.gam * acc[v][1]);
        const __amdgpu_buffer_rsrc_t rx = rs(a.xout, v, j);
        stm(rx, o0, a.pu * N[v][0] + a.px * Xo[v][0] + a.pin * Uo[v][0] + a.pl * acc[v][0]);
        stm(rx, o1, a.pu * N[v][1] + a.px * Xo[v][1] + a.pin * Uo[v][1] + a.pl * acc[v][1]);
      }
    }
  };

  // ---- prologue: top traces of the row below the segment, first row's modes ------
  float B0[4][2], B1[4][2], N[4][2];
  Vec4<float> Tt = {0.f, 0.f, 0.f, 0.f};   // top trace (point gq) of the row below the current one
  load_row(a.in, y0, B0);
  if (y0 > 0) {
    float Bh[4][2];
    load_row(a.in, y0 - 1, Bh);
    v4f df[4];
    fwd_face(Bh, df);
    Tt = {df[0][3], df[1][3], df[2][3], df[3][3]};
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) N[v][0] = N[v][1] = 0.f;
  v4f accA[4], accB[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) accA[v] = accB[v] = v4f{0.f, 0.f, 0.f, 0.f};

  // one row: Bc holds its modes, Bn receives the next row's, accp holds the
  // accumulator of the row below (finished here), acc receives this row's
  auto row = [&](int j, const float (&Bc)[4][2], float (&Bn)[4][2], v4f (&accp)[4], v4f (&acc)[4]) {
    // loads in flight during this row: the next row, and U^n of the row finished
    // here.  They are issued on every row (without U^n, through offsets that
    // read 0), and the first row's finish only drops its stores, so the body of
    // a row pair has no branch and the scheduler can overlap one row's products
    // with the next row's.
    load_row(a.in, j + 1, Bn);
    load_own_if(with_un, un_base, j - 1, N);
    if constexpr (DUAL) {
      load_own(a.in, j - 1, Uo);
      load_own(a.xin, j - 1, Xo);
    }

    v4f dn[4], df[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const v4f z = {0.f, 0.f, 0.f, 0.f};
      dn[v] = mfma(An0, Bc[v][0], z);
      dn[v] = mfma(An1, Bc[v][1], dn[v]);
    }
    fwd_face(Bc, df);
    // accumulator start: U plus dt times the modal source -g (0, rho, rho, mx + my)
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      acc[0][s] = ku * Bc[0][s];
      acc[1][s] = ku * Bc[1][s] - gdt * Bc[0][s];
      acc[2][s] = ku * Bc[2][s] - gdt * Bc[0][s];
      acc[3][s] = ku * Bc[3][s] - gdt * (Bc[1][s] + Bc[2][s]);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[v][2] = acc[v][3] = 0.f;

    // volume: fluxes at this group's node of each step, fed back as k-slot g.
    // The three y nodes' factors are scalar loads, then a per-lane choice (a
    // choice between the loads themselves becomes a branch around them).
    const double ey0 = ctab(a.eyn, 3 * j), ey1 = ctab(a.eyn, 3 * j + 1), ey2 = ctab(a.eyn, 3 * j + 2);
    const double eyg = g == 0 ? ey0 : (g == 1 ? ey1 : ey2);
    float X2[4], Y2[4];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      const double eyt = nyb[t] == 0 ? ey0 : (nyb[t] == 1 ? ey1 : ey2);
      const FaceEq<float> q = eq_at(a, exr[t] * eyt);
      Vec4<float> fx, fy;
      point_terms(P, q, Vec4<float>{dn[0][t], dn[1][t], dn[2][t], dn[3][t]}, fx, fy);
      const float X[4] = {fx.r, fx.mx, fx.my, fx.e};
      const float Y[4] = {fy.r, fy.mx, fy.my, fy.e};
      if (t < 2) {
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[v] = mfma(Ab[2 * t], X[v], acc[v]);
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[v] = mfma(Ab[2 * t + 1], Y[v], acc[v]);
      } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) { X2[v] = X[v]; Y2[v] = Y[v]; }
      }
    }

    // x face on the left of column i at point gq: right trace of column i-1 | left trace of i
    Vec4<float> xl = {from_col_left(df[0][1]), from_col_left(df[1][1]), from_col_left(df[2][1]),
                      from_col_left(df[3][1])};
    Vec4<float> xr = {df[0][0], df[1][0], df[2][0], df[3][0]};
    if (i == 0) xl = mirror_x(xr);
    if (i == nx) xr = mirror_x(xl);
    const Vec4<float> Fx = face_flux<float, FLUX, 0, FORM_DEV>(P, eq_at(a, exf_i * eyg), xl, xr);
    const float fl[4] = {Fx.r, Fx.mx, Fx.my, Fx.e};
    // y face below row j at point gq: top trace of row j-1 | bottom trace of row j
    const Vec4<float> yr = {df[0][2], df[1][2], df[2][2], df[3][2]};
    const Vec4<float> Fb = flux_y(j, j == 0 ? mirror_y(yr) : Tt, yr);
    const float fb[4] = {Fb.r, Fb.mx, Fb.my, Fb.e};
    const bool g3 = g == 3;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const float fr = from_col_right(fl[v]);   // right face = left face of column i+1
      acc[v] = mfma(Ab[4], g3 ? X2[v] : fl[v], acc[v]);
      acc[v] = mfma(Ab[5], g3 ? Y2[v] : fr, acc[v]);
      acc[v] = mfma(Ab[6], fb[v], acc[v]);
    }
    // the face below row j is the top face of row j-1
    finish(j - 1, j > y0, accp, Fb, N);
    Tt = {df[0][3], df[1][3], df[2][3], df[3][3]};
  };

  // ---- epilogue: the top face of the segment's last row (Bc holds row y1) ----------
  auto last = [&](const float (&Bc)[4][2], v4f (&accp)[4]) {
    if (with_un) load_own(a.un, y1 - 1, N);
    if (DUAL) {
      load_own(a.in, y1 - 1, Uo);
      load_own(a.xin, y1 - 1, Xo);
    }
    Vec4<float> yr;
    if (y1 < ny) {
      v4f df[4];
      fwd_face(Bc, df);
      yr = {df[0][2], df[1][2], df[2][2], df[3][2]};
    } else {
      yr = mirror_y(Tt);
    }
    finish(y1 - 1, true, accp, flux_y(y1, Tt, yr), N);
  };

  // rows in pairs with the register sets swapped, so nothing is copied between rows
  int j = y0;
  for (; j + 1 < y1; j += 2) {
    row(j, B0, B1, accA, accB);
    row(j + 1, B1, B0, accB, accA);
    // the row pair is one basic block: ask the scheduler to alternate two
    // matrix products with 16 VALU instructions through it, so the products'
    // latency is covered by the point-wise flux work (1.3 % faster than the
    // default order; 1:4, 1:6, 1:8 and 1:12 measured the same within noise)
#pragma unroll
    for (int q = 0; q < 48; ++q) {
      __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);    // MFMA
      __builtin_amdgcn_sched_group_barrier(0x002, 16, 0);   // VALU
    }
  }
  if (j < y1) {
    row(j, B0, B1, accA, accB);
    last(B1, accB);
  } else {
    last(B0, accA);
  }
}

// TVB limiter (in place; means are never written, so neighbours may be read
// concurrently) and, when requested, the CFL rate of the limited state plus the
// step bookkeeping of the instance.  One thread per cell, a row of cells per
// grid row; every mode of variable v is read through one buffer resource, so
// out-of-range neighbours and masked lanes cost no address arithmetic.
struct LimArgs {
  float* st;           // cell (0,0) of the 24 mode planes
  const double* exn;   // e^{-k x} at x nodes [3*nx] (FP64)
  const double* eyn;   // e^{-k y} at y nodes [3*ny]
  Ctrl* ctrl;
  long long plane;
  int pitch, nx, ny, parity;
  int nrec;            // bytes from a variable's mode-0 cell (0,0) to the end of its six planes
  int limit, charlim, posit, cfl, bookkeeping, init;
  const double* eqa;   // A_n(x cell) [3*nx] and B_n(y cell) [3*ny] (FP64): equilibrium
  const double* eqb;   //   modes are rho0 A_a B_b (density) and eover A_a B_b (energy)
  double t_end, cfl_over;
  double rho0, eover;  // equilibrium rho and E per e^{-k(x+y)}
  float gamma, odx, ody, mdx2, mdy2;
  float phi1[3], phi2[3];   // phi_1, phi_2 at the Gauss nodes
};

__device__ inline float gminmod(float x, float y, float z, float mh2) {
  if (fabsf(x) < mh2) return x;
  const bool sx = signbit(x);
  if (signbit(y) != sx || signbit(z) != sx) return 0.f;
  const float m = fminf(fminf(fabsf(x), fabsf(y)), fabsf(z));
  return sx ? -m : m;
}

// Eigenvectors of the Euler flux Jacobian along x (DIR 0) or y at the state u:
// columns of R for the speeds vn - c, vn, vn (shear), vn + c; L = R^-1.
template <int DIR>
__device__ inline void euler_eigen(float gamma, const float u[4], float R[4][4], float L[4][4]) {
  const float irho = 1.0f / u[0];
  const float vx = u[1] * irho, vy = u[2] * irho, q2 = vx * vx + vy * vy;
  const float p = (gamma - 1.0f) * (u[3] - 0.5f * u[0] * q2);
  const float c = sqrtf(gamma * fmaxf(p, 1e-10f) * irho), ic = 1.0f / c;
  const float H = (u[3] + p) * irho;
  const float vn = DIR == 0 ? vx : vy, vt = DIR == 0 ? vy : vx;
  constexpr int n = DIR == 0 ? 1 : 2, t = DIR == 0 ? 2 : 1;
  const float b1 = (gamma - 1.0f) * ic * ic, b2 = 0.5f * b1 * q2;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int k = 0; k < 4; ++k) R[i][k] = L[i][k] = 0.f;
  R[0][0] = 1.f;         R[0][1] = 1.f;       R[0][3] = 1.f;
  R[n][0] = vn - c;      R[n][1] = vn;        R[n][3] = vn + c;
  R[t][0] = vt;          R[t][1] = vt;        R[t][2] = 1.f;  R[t][3] = vt;
  R[3][0] = H - vn * c;  R[3][1] = 0.5f * q2; R[3][2] = vt;   R[3][3] = H + vn * c;
  L[0][0] = 0.5f * (b2 + vn * ic); L[0][n] = -0.5f * (b1 * vn + ic); L[0][t] = -0.5f * b1 * vt; L[0][3] = 0.5f * b1;
  L[1][0] = 1.f - b2;              L[1][n] = b1 * vn;                L[1][t] = b1 * vt;         L[1][3] = -b1;
  L[2][0] = -vt;                                                     L[2][t] = 1.f;
  L[3][0] = 0.5f * (b2 - vn * ic); L[3][n] = -0.5f * (b1 * vn - ic); L[3][t] = -0.5f * b1 * vt; L[3][3] = 0.5f * b1;
}

// solve_for_t (2d/limiters.f90:312-362) in FP64: the t in [0,1] at which the
// pressure of (1-t) avg + t u reaches eps
__device__ inline float solve_for_t(double gamma, double eps, const float uf[4], const float af[4]) {
  const double pa = af[0], mxa = af[1], mya = af[2], ea = af[3];
  const double pj = uf[0], mxj = uf[1], myj = uf[2], ej = uf[3];
  double qa = 2.0 * (pj - pa) * (ej - ea) - (mxj - mxa) * (mxj - mxa) - (myj - mya) * (myj - mya);
  double qb = 2.0 * (pj - pa) * (ea - eps / (gamma - 1)) + 2.0 * pa * (ej - ea) -
              2.0 * (mxa * (mxj - mxa) + mya * (myj - mya));
  double qc = 2.0 * pa * ea - (mxa * mxa + mya * mya) - 2.0 * eps * pa / (gamma - 1.0);
  qb = qb / qa;
  qc = qc / qa;
  const double D = sqrt(fabs(qb * qb - 4 * qc));
  const double t1 = 0.5 * (-qb - D), t2 = 0.5 * (-qb + D);
  double t = 0.0;
  if (t1 > -eps && t1 < 1.0 + eps) t = t1;
  else if (t2 > -eps && t2 < 1.0 + eps) t = t2;
  return (float)fmax(0.0, fmin(1.0, t));
}

// Positivity limiter (compute_positivity, 2d/limiters.f90:478-654) on the full
// state F = E + md of one cell: E the equilibrium's modes from the separable
// tables; the set is {GLL x GL} u {GL x GLL} u {GL x GL}; thresholds 1e-5 of the
// cell's equilibrium density and pressure.  Marks each variable it rewrites.
__device__ inline void positivity_cell(float md[4][NM], bool chg[4], const LimArgs& a, int i, int j) {
  const double ax[3] = {a.eqa[3 * i], a.eqa[3 * i + 1], a.eqa[3 * i + 2]};
  const double by[3] = {a.eqb[3 * j], a.eqb[3 * j + 1], a.eqb[3 * j + 2]};
  constexpr int MI_[NM] = {0, 1, 0, 2, 1, 0}, MJ_[NM] = {0, 0, 1, 0, 1, 2};
  const double cv[4] = {a.rho0, 0.0, 0.0, a.eover};
  float E[4][NM], F[4][NM];
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      E[v][m] = (float)(cv[v] * (ax[MI_[m]] * by[MJ_[m]]));   // formed in FP64, rounded once
      F[v][m] = E[v][m] + md[v][m];
    }
  const float gm1 = a.gamma - 1.0f;
  const float eps_r = fmaxf(1e-10f, 1e-5f * E[0][0]);
  const float eps_p = fmaxf(1e-10f, 1e-5f * gm1 * E[3][0]);
  // phi_1, phi_2 at the Gauss-Lobatto points -1, 0, 1 and at the Gauss points
  const float s3 = 1.7320508075688772f, s5 = 2.2360679774997896f;
  const float g1[3] = {-s3, 0.f, s3}, g2[3] = {s5, -0.5f * s5, s5};
  auto val = [&](int v, float p1x, float p2x, float p1y, float p2y) {
    return F[v][0] + F[v][1] * p1x + F[v][2] * p1y + F[v][3] * p2x + F[v][4] * (p1x * p1y) + F[v][5] * p2y;
  };
  // set point k in 0..26: q = k/9, c = (k/3)%3, side = k%3
  auto point = [&](int k, float& p1x, float& p2x, float& p1y, float& p2y) {
    const int q = k / 9, c = (k / 3) % 3, side = k % 3;
    if (side == 0) { p1x = g1[c]; p2x = g2[c]; p1y = a.phi1[q]; p2y = a.phi2[q]; }
    else if (side == 1) { p1x = a.phi1[q]; p2x = a.phi2[q]; p1y = g1[c]; p2y = g2[c]; }
    else { p1x = a.phi1[q]; p2x = a.phi2[q]; p1y = a.phi1[c]; p2y = a.phi2[c]; }
  };
  float rmin = INFINITY;
  for (int k = 0; k < 27; ++k) {
    float p1x, p2x, p1y, p2y;
    point(k, p1x, p2x, p1y, p2y);
    rmin = fminf(rmin, val(0, p1x, p2x, p1y, p2y));
  }
  const float theta = fminf(fabsf((F[0][0] - eps_r) / (F[0][0] - rmin)), 1.f);
  if (theta < 1.f) {
#pragma unroll
    for (int m = 1; m < NM; ++m) {
      F[0][m] = theta * F[0][m];
      md[0][m] = F[0][m] - E[0][m];
    }
    chg[0] = true;
  }
  const float avg[4] = {F[0][0], F[1][0], F[2][0], F[3][0]};
  float tmin = 1.f;
  for (int k = 0; k < 27; ++k) {
    float p1x, p2x, p1y, p2y;
    point(k, p1x, p2x, p1y, p2y);
    float u[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) u[v] = val(v, p1x, p2x, p1y, p2y);
    const float r = fmaxf(u[0], 1e-9f);
    const float vx = u[1] / r, vy = u[2] / r;
    const float pr = gm1 * (u[3] - 0.5f * r * (vx * vx + vy * vy));
    if (!(pr > eps_p)) tmin = fminf(tmin, solve_for_t(a.gamma, eps_p, u, avg));
  }
  if (tmin < 1.f) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
#pragma unroll
      for (int m = 1; m < NM; ++m) md[v][m] = tmin * F[v][m] - E[v][m];
      chg[v] = true;
    }
  }
}

// TVB, CHAR and POS select the limiters compiled in (each launch carries only
// its own code); the CFL rate and the bookkeeping stay runtime switches.
template <bool TVB, bool CHAR, bool POS>
__global__ __launch_bounds__(256) void dg_limit_kernel(LimArgs a) {
  const int par = a.parity;
  const int blin = blockIdx.y * gridDim.x + blockIdx.x;
  const bool active = a.init || (a.ctrl->t[par] < a.t_end);
  if (!active) {
    // instance done: carry the clock and rate to the next parity once
    if (a.bookkeeping && blin == 0 && threadIdx.x < RSLOTS) {
      if (threadIdx.x == 0) a.ctrl->t[par ^ 1] = a.ctrl->t[par];
      a.ctrl->rate[par ^ 1][threadIdx.x].v = a.ctrl->rate[par][threadIdx.x].v;
    }
    return;
  }
  const int nx = a.nx, ny = a.ny;
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  const bool in = i < nx;
  const unsigned pb = (unsigned)a.plane * 4u;
  const unsigned ob = (unsigned)(j * a.pitch + i) * 4u;
  auto off = [&](bool ok, unsigned o, int m) { return ok ? o + (unsigned)m * pb : MASKED; };
  __amdgpu_buffer_rsrc_t rv[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) rv[v] = mode_rsrc(a.st + (long long)(v * NM) * a.plane, a.nrec);
  // TVB needs modes 0-2 only; the CFL rate and the positivity limiter need all six
  float md[4][NM];
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int m = 0; m < NM; ++m) md[v][m] = (m < 3 || a.cfl || POS) ? ldm(rv[v], off(in, ob, m)) : 0.f;
  bool chg[4] = {false, false, false, false};

  if (TVB) {
    const float s3 = 1.7320508075688772f;
    const bool hw = i > 0, he = i < nx - 1, hs = j > 0, hn = j < ny - 1;
    const unsigned ow = off(in && hw, ob - 4u, 0), oe = off(in && he, ob + 4u, 0);
    const unsigned os = off(in && hs, ob - (unsigned)a.pitch * 4u, 0), on = off(in && hn, ob + (unsigned)a.pitch * 4u, 0);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const float u0 = md[v][0];
      // walls mirror the mean, with the normal momentum negated
      const float mx = v == 1 ? -u0 : u0, my = v == 2 ? -u0 : u0;
      const float lw = ldm(rv[v], ow), le = ldm(rv[v], oe), ls = ldm(rv[v], os), ln = ldm(rv[v], on);
      const float uw = hw ? lw : mx, ue = he ? le : mx, us = hs ? ls : my, un = hn ? ln : my;
      const float x = s3 * md[v][1], y = s3 * md[v][2];
      const float xt = gminmod(x, ue - u0, u0 - uw, a.mdx2);
      const float yt = gminmod(y, un - u0, u0 - us, a.mdy2);
      const bool lim = in && (xt != x || yt != y);
      if (lim) {
        md[v][1] = xt * (1.0f / 1.7320508075688772f);
        md[v][2] = yt * (1.0f / 1.7320508075688772f);
        md[v][3] = md[v][4] = md[v][5] = 0.f;
        chg[v] = true;
      }
    }
  }
  if (CHAR && in) {
    // characteristic TVB: the neighbour means' differences and the linear modes
    // in the characteristic variables of the cell's full-state mean
    const float s3 = 1.7320508075688772f;
    const bool hw = i > 0, he = i < nx - 1, hs = j > 0, hn = j < ny - 1;
    const unsigned ow = off(in && hw, ob - 4u, 0), oe = off(in && he, ob + 4u, 0);
    const unsigned os = off(in && hs, ob - (unsigned)a.pitch * 4u, 0), on = off(in && hn, ob + (unsigned)a.pitch * 4u, 0);
    float nbw[4], nbe[4], nbs[4], nbn[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const float u0 = md[v][0];
      const float mx = v == 1 ? -u0 : u0, my = v == 2 ? -u0 : u0;
      const float lw = ldm(rv[v], ow), le = ldm(rv[v], oe), ls = ldm(rv[v], os), ln = ldm(rv[v], on);
      nbw[v] = hw ? lw : mx; nbe[v] = he ? le : mx; nbs[v] = hs ? ls : my; nbn[v] = hn ? ln : my;
    }
    const double e00 = a.eqa[3 * i] * a.eqb[3 * j];
    const float ubar[4] = {(float)(a.rho0 * e00) + md[0][0], md[1][0], md[2][0],
                           (float)(a.eover * e00) + md[3][0]};
    float lin[2][4];
    bool changed = false;
    auto limit_dir = [&](auto dirc, const float* up, const float* um, float mh2) {
      constexpr int D = decltype(dirc)::value;
      float R[4][4], L[4][4], av[4], dp[4], dm[4];
      euler_eigen<D>(a.gamma, ubar, R, L);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        av[v] = s3 * md[v][1 + D];
        dp[v] = up[v] - md[v][0];
        dm[v] = md[v][0] - um[v];
      }
      float cl[4];
      bool lim = false;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        float ca = 0.f, cp = 0.f, cm = 0.f;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          ca += L[k][v] * av[v];
          cp += L[k][v] * dp[v];
          cm += L[k][v] * dm[v];
        }
        cl[k] = gminmod(ca, cp, cm, mh2);
        lim |= cl[k] != ca;
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < 4; ++k) s += R[v][k] * cl[k];
        lin[D][v] = lim ? s * (1.0f / 1.7320508075688772f) : md[v][1 + D];
      }
      changed |= lim;
    };
    limit_dir(std::integral_constant<int, 0>{}, nbe, nbw, a.mdx2);
    limit_dir(std::integral_constant<int, 1>{}, nbn, nbs, a.mdy2);
    if (changed) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        md[v][1] = lin[0][v];
        md[v][2] = lin[1][v];
        md[v][3] = md[v][4] = md[v][5] = 0.f;
        chg[v] = true;
      }
    }
  }
  if (POS && in) positivity_cell(md, chg, a, i, j);
  if (TVB || CHAR || POS) {
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int m = 1; m < NM; ++m) stm(rv[v], off(chg[v], ob, m), md[v][m]);
  }

  float rate = 0.f;
  bool bad = false;
  if (a.cfl && in) {
    const float gm1 = a.gamma - 1.0f;
    const int ic = i;
    double ex[NQ], ey[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) { ex[q] = a.exn[3 * ic + q]; ey[q] = a.eyn[3 * j + q]; }
    // node values by sum factorisation: d(x_a, y_b) = A0(a) + phi_1(y_b) A1(a) + phi_2(y_b) m5,
    // A0 = m0 + phi_1 m1 + phi_2 m3, A1 = m2 + phi_1 m4 (modes (0,0),(1,0),(0,1),(2,0),(1,1),(0,2))
    float A0[4][NQ], A1[4][NQ];
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        A0[v][q] = md[v][0] + a.phi1[q] * md[v][1] + a.phi2[q] * md[v][3];
        A1[v][q] = md[v][2] + a.phi1[q] * md[v][4];
      }
#pragma unroll
    for (int qb = 0; qb < NQ; ++qb)
#pragma unroll
      for (int qa = 0; qa < NQ; ++qa) {
        const double e = ex[qa] * ey[qb];
        float d[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) d[v] = A0[v][qa] + a.phi1[qb] * A1[v][qa] + a.phi2[qb] * md[v][5];
        const float rho = (float)(a.rho0 * e) + d[0];
        const float inv = trcp(fmaxf(rho, 1e-10f));
        const float u = d[1] * inv, w = d[2] * inv;
        const float p = gm1 * (((float)(a.eover * e) + d[3]) - 0.5f * rho * (u * u + w * w));
        const float cs = tsqrt(a.gamma * fmaxf(p, 1e-10f) * inv);
        const float rt = (fabsf(u) + cs) * a.odx + (fabsf(w) + cs) * a.ody;
        bad |= !(rho > 0.f) || !(p > 0.f) || !(rt < 1e30f);
        rate = fmaxf(rate, rt);
      }
  }
  if (a.cfl) {
    // workgroup maximum, then one atomic per workgroup into its slot
    __shared__ float wmax[4];
    for (int o = 32; o > 0; o >>= 1) rate = fmaxf(rate, __shfl_xor(rate, o));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = rate;
    const int anybad = __syncthreads_or(bad);
    const int slot = a.init ? par : (par ^ 1);
    if (threadIdx.x == 0) {
      const float m = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
      atomicMax(&a.ctrl->rate[slot][blin % RSLOTS].v, __float_as_uint(m));
      if (anybad) atomicOr(&a.ctrl->err, 1u);
    }
    if (a.bookkeeping && blin == 0 && threadIdx.x == 0) {
      // the step that just ended: same dt as its stages computed
      const double t = a.ctrl->t[par];
      double dt = a.cfl_over / (double)ctrl_rate(a.ctrl, par);
      if (t + dt > a.t_end) dt = a.t_end - t;
      a.ctrl->t[par ^ 1] = t + dt;
      a.ctrl->dt_last = dt;
      a.ctrl->steps += 1;
    }
  }
}

// host AoS(4,nx,ny,6) FP64 <-> device planes FP32
__global__ void dg_to_planes_kernel(const double* __restrict__ h, float* __restrict__ st, int nx, int ny, int pitch,
                                    long long plane) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = 4LL * nx * ny * NM;
  if (q >= n) return;
  const int v = (int)(q % 4);
  const long long rest = q / 4;
  const int i = (int)(rest % nx), j = (int)((rest / nx) % ny), m = (int)(rest / ((long long)nx * ny));
  st[(long long)(v * NM + m) * plane + (long long)j * pitch + i] = (float)h[q];
}

__global__ void dg_from_planes_kernel(const float* __restrict__ st, double* __restrict__ h, int nx, int ny,
                                      int pitch, long long plane) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = 4LL * nx * ny * NM;
  if (q >= n) return;
  const int v = (int)(q % 4);
  const long long rest = q / 4;
  const int i = (int)(rest % nx), j = (int)((rest / nx) % ny), m = (int)(rest / ((long long)nx * ny));
  h[q] = (double)st[(long long)(v * NM + m) * plane + (long long)j * pitch + i];
}

__global__ void dg_fill_kernel(float* __restrict__ p, long long n) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) p[q] = 0.f;
}

}  // namespace dg
}  // namespace fvwb

using namespace fvwb::dg;

// ---- basis on the host (same functions as 2d/legendre.f90, FP64) --------------------------
static double leg(double x, int n) {
  const double l = n == 0 ? 1.0 : (n == 1 ? x : 0.5 * (3 * x * x - 1));
  return std::sqrt(2.0 * n + 1.0) * l;
}
static double legp(double x, int n) {
  const double l = n == 0 ? 0.0 : (n == 1 ? 1.0 : 3.0 * x);
  return std::sqrt(2.0 * n + 1.0) * l;
}
static const int MI[6] = {0, 1, 0, 2, 1, 0}, MJ[6] = {0, 0, 1, 0, 1, 2};

struct fvwb_dg {
  fvwb_dg_params p{};
  hipStream_t stream = nullptr;
  int pitch = 0;
  long long plane = 0;
  long long org = 0;              // offset of cell (0,0) inside the ghost frame
  int ntx = 0, nseg = 0, seg_rows = 0;
  float* buf[4] = {nullptr, nullptr, nullptr, nullptr};   // SSP(5,4) uses the fourth
  int nbuf = 3;
  double* host_stage = nullptr;   // device FP64 AoS staging
  double *exn = nullptr, *eyn = nullptr, *exf = nullptr, *eyf = nullptr;
  double *eqa = nullptr, *eqb = nullptr;   // 1D Gauss projections of e^{-kx}, e^{-ky} per cell
  float* cell(int b) const { return buf[b] + org; }
  Consts* C = nullptr;
  Ctrl* ctrl = nullptr;
  int parity = 0;
  bool ready = false;
  double xq[3], wq[3];
  std::string err;
};

static thread_local std::string g_dg_err;
static int dg_fail(fvwb_dg* h, int code, const std::string& m) {
  if (h) h->err = m; else g_dg_err = m;
  return code;
}
#define DGHIP(h, expr)                                                                                   \
  do {                                                                                                   \
    hipError_t e_ = (expr);                                                                              \
    if (e_ != hipSuccess) return dg_fail((h), FVWB_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

extern "C" const char* fvwb_dg_last_error(fvwb_dg_handle h) { return h ? h->err.c_str() : g_dg_err.c_str(); }

extern "C" int fvwb_dg_destroy(fvwb_dg_handle h) {
  if (!h) return FVWB_OK;
  void* ps[] = {h->buf[0], h->buf[1], h->buf[2], h->buf[3], h->host_stage, h->exn, h->eyn, h->exf, h->eyf,
                h->eqa, h->eqb, h->C, h->ctrl};
  for (void* q : ps)
    if (q) (void)hipFree(q);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return FVWB_OK;
}

extern "C" int fvwb_dg_create(fvwb_dg_handle* out, const fvwb_dg_params* p) {
  if (!out) return dg_fail(nullptr, FVWB_ERR_ARG, "handle pointer is NULL");
  *out = nullptr;
  if (!p) return dg_fail(nullptr, FVWB_ERR_ARG, "params is NULL");
  if (p->nvar != 4 || p->nx < 2 || p->ny < 2) return dg_fail(nullptr, FVWB_ERR_ARG, "nvar must be 4, nx, ny >= 2");
  if (p->degree != 2) return dg_fail(nullptr, FVWB_ERR_ARG, "the device path implements degree 2 (P2)");
  if (p->flux != FVWB_FLUX_RUSANOV && p->flux != FVWB_FLUX_HLL && p->flux != FVWB_FLUX_HLLC)
    return dg_fail(nullptr, FVWB_ERR_ARG, "bad flux");
  if (p->limiter < FVWB_DG_LIMITER_NONE || p->limiter > FVWB_DG_LIMITER_CHARACTERISTIC_POSITIVITY)
    return dg_fail(nullptr, FVWB_ERR_ARG, "bad limiter");
  if (p->bc != FVWB_BC_REFLECT) return dg_fail(nullptr, FVWB_ERR_ARG, "the DG path implements walls (FVWB_BC_REFLECT)");
  if (p->integrator != FVWB_DG_SSPRK3 && p->integrator != FVWB_DG_SSPRK54)
    return dg_fail(nullptr, FVWB_ERR_ARG, "integrator must be FVWB_DG_SSPRK3 or FVWB_DG_SSPRK54");
  if (!(p->gamma > 1.0) || !(p->boxlen_x > 0) || !(p->boxlen_y > 0) || !(p->cfl > 0) || !(p->rho0 > 0) ||
      !(p->p0 > 0) || !(p->g >= 0) || !(p->tvb_m >= 0))
    return dg_fail(nullptr, FVWB_ERR_ARG, "gamma > 1, positive box, cfl, rho0, p0 and g, M >= 0 required");
  if ((long long)p->nx * p->ny > (1LL << 26)) return dg_fail(nullptr, FVWB_ERR_ARG, "nx*ny must be <= 2^26");
  fvwb_dg* h = new (std::nothrow) fvwb_dg();
  if (!h) return dg_fail(nullptr, FVWB_ERR_OOM, "host allocation failed");
  h->p = *p;
  const int nx = p->nx, ny = p->ny;
  // ghost frame: one column left of column 0 and enough on the right for the
  // last strip's lanes, one row below row 0 and one above row ny-1; the ghost
  // cells are zero for the life of the handle and never written
  h->ntx = (nx + SW - 1) / SW;
  h->pitch = (SW * h->ntx + 2 + 15) / 16 * 16;
  h->plane = ((long long)h->pitch * (ny + GROWS) + 15) / 16 * 16;
  h->org = h->pitch + 1;
  if (NM * h->plane * 4 >= (1LL << 31)) {
    delete h;
    return dg_fail(nullptr, FVWB_ERR_ARG, "grid too large for the 2 GiB mode-plane group of one variable");
  }
  auto bail = [&](int rc, const char* m) { h->err = m; g_dg_err = m; fvwb_dg_destroy(h); return rc; };
  if (hipStreamCreate(&h->stream) != hipSuccess) return bail(FVWB_ERR_HIP, "hipStreamCreate failed");
  const size_t sbytes = sizeof(float) * 4 * NM * (size_t)h->plane;
  h->nbuf = p->integrator == FVWB_DG_SSPRK54 ? 4 : 3;
  for (int b = 0; b < h->nbuf; ++b)
    if (hipMalloc(&h->buf[b], sbytes) != hipSuccess) return bail(FVWB_ERR_OOM, "device allocation failed");
  if (hipMalloc(&h->host_stage, sizeof(double) * 4 * NM * (size_t)nx * ny) != hipSuccess ||
      hipMalloc(&h->exn, sizeof(double) * 3 * nx) != hipSuccess || hipMalloc(&h->eyn, sizeof(double) * 3 * ny) != hipSuccess ||
      hipMalloc(&h->exf, sizeof(double) * (nx + 1)) != hipSuccess ||
      hipMalloc(&h->eyf, sizeof(double) * (ny + 1)) != hipSuccess || hipMalloc(&h->C, sizeof(Consts)) != hipSuccess ||
      hipMalloc(&h->eqa, sizeof(double) * 3 * nx) != hipSuccess || hipMalloc(&h->eqb, sizeof(double) * 3 * ny) != hipSuccess ||
      hipMalloc(&h->ctrl, sizeof(Ctrl)) != hipSuccess)
    return bail(FVWB_ERR_OOM, "device allocation failed");
  // padding columns and rows stay zero for the whole life of the handle
  for (int b = 0; b < h->nbuf; ++b) {
    const long long n = 4LL * NM * h->plane;
    hipLaunchKernelGGL(dg_fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, h->buf[b], n);
  }
  // Gauss-Legendre 3-point rule: the nodes are -+sqrt(3/5), 0 (the reference's
  // gl_quadrature(3) converges to these to the last bit)
  const double s35 = std::sqrt(0.6);
  h->xq[0] = -s35; h->xq[1] = 0.0; h->xq[2] = s35;
  h->wq[0] = 5.0 / 9.0; h->wq[1] = 8.0 / 9.0; h->wq[2] = 5.0 / 9.0;
  const double dx = p->boxlen_x / nx, dy = p->boxlen_y / ny, kk = p->rho0 * p->g / p->p0;
  std::vector<double> ex(3 * nx), ey(3 * ny), fx(nx + 1), fy(ny + 1);
  for (int i = 0; i < nx; ++i)
    for (int a = 0; a < 3; ++a) ex[3 * i + a] = std::exp(-(kk * ((i + 0.5) * dx + 0.5 * dx * h->xq[a])));
  for (int j = 0; j < ny; ++j)
    for (int b = 0; b < 3; ++b) ey[3 * j + b] = std::exp(-(kk * ((j + 0.5) * dy + 0.5 * dy * h->xq[b])));
  for (int i = 0; i <= nx; ++i) fx[i] = std::exp(-(kk * (i * dx)));
  for (int j = 0; j <= ny; ++j) fy[j] = std::exp(-(kk * (j * dy)));
  // the equilibrium's modes for the positivity limiter: U_eq is separable, so
  // its Gauss projection on mode (a, b) is c_v A_a(x cell) B_b(y cell) with
  // A_a = 1/2 sum_q w_q e^{-k x_q} phi_a(xi_q)
  std::vector<double> eqa(3 * nx), eqb(3 * ny);
  for (int i = 0; i < nx; ++i)
    for (int n = 0; n < 3; ++n) {
      double s = 0.0;
      for (int q = 0; q < 3; ++q) s += 0.5 * h->wq[q] * ex[3 * i + q] * leg(h->xq[q], n);
      eqa[3 * i + n] = s;
    }
  for (int j = 0; j < ny; ++j)
    for (int n = 0; n < 3; ++n) {
      double s = 0.0;
      for (int q = 0; q < 3; ++q) s += 0.5 * h->wq[q] * ey[3 * j + q] * leg(h->xq[q], n);
      eqb[3 * j + n] = s;
    }
  Consts C;
  std::memset(&C, 0, sizeof C);
  const double* xq = h->xq;
  const double* wq = h->wq;
  // volume weights of node n: 0.5 w_a w_b dphi/dx phi and phi dphi/dy there
  auto wvol = [&](int n, int m, bool xdir) {
    const int na = n % 3, nb = n / 3;
    const double w = wq[na] * wq[nb];
    return xdir ? 0.5 * w * legp(xq[na], MI[m]) * leg(xq[nb], MJ[m]) : 0.5 * w * leg(xq[na], MI[m]) * legp(xq[nb], MJ[m]);
  };
  for (int row = 0; row < 16; ++row) {
    const int g = row >> 2, r = row & 3;
    for (int m = 0; m < NM; ++m) {
      if ((g < 3 && r < 2) || (g == 3 && r < 3)) {
        const int n = vol_node(g, r);
        C.fwd_node[row][m] = (float)(leg(xq[n % 3], MI[m]) * leg(xq[n / 3], MJ[m]));
      }
      if (g < 3) {
        const int q = g;
        double v;
        if (r == 0) v = leg(-1.0, MI[m]) * leg(xq[q], MJ[m]);        // left trace
        else if (r == 1) v = leg(1.0, MI[m]) * leg(xq[q], MJ[m]);    // right
        else if (r == 2) v = leg(xq[q], MI[m]) * leg(-1.0, MJ[m]);   // bottom
        else v = leg(xq[q], MI[m]) * leg(1.0, MJ[m]);                // top
        C.fwd_face[row][m] = (float)v;
      }
    }
    // output row -> mode: row 4g is mode g, row 4g+1 mode g+4 (g < 2), others none
    const int m = r == 0 ? g : (r == 1 && g < 2 ? g + 4 : -1);
    if (m < 0) continue;
    for (int k = 0; k < 4; ++k) {          // k = the lane group whose value takes the slot
      // steps 0 and 1: F_x and F_y of the group's node
      for (int t = 0; t < 2; ++t) {
        C.back[2 * t][row][k] = (float)wvol(vol_node(k, t), m, true);
        C.back[2 * t + 1][row][k] = (float)wvol(vol_node(k, t), m, false);
      }
      if (k < 3) {
        // face fluxes (pre-scaled by 1/dx or 1/dy) at point k: +-0.5 w phi at the face
        C.back[4][row][k] = (float)(0.5 * wq[k] * leg(-1.0, MI[m]) * leg(xq[k], MJ[m]));
        C.back[5][row][k] = (float)(-0.5 * wq[k] * leg(1.0, MI[m]) * leg(xq[k], MJ[m]));
        C.back[6][row][k] = (float)(0.5 * wq[k] * leg(xq[k], MI[m]) * leg(-1.0, MJ[m]));
        C.back_top[row][k] = (float)(-0.5 * wq[k] * leg(xq[k], MI[m]) * leg(1.0, MJ[m]));
      } else {
        // group 3's third node
        C.back[4][row][k] = (float)wvol(vol_node(3, 2), m, true);
        C.back[5][row][k] = (float)wvol(vol_node(3, 2), m, false);
      }
    }
  }
  // segments along y: as many waves as the device holds at once, at least 4 rows each
  {
    int dev = 0, cus = 0, blocks_per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks_per_cu, dg_stage_kernel<fvwb::FLUX_HLL, false>, 256, 0) !=
            hipSuccess)
      return bail(FVWB_ERR_HIP, "device query failed");
    const long long slots = 4LL * std::max(1, cus) * std::max(1, blocks_per_cu);
    const long long nseg = std::max(1LL, std::min<long long>(ny, slots / h->ntx));
    h->seg_rows = std::max(4, (int)((ny + nseg - 1) / nseg));
    h->nseg = (ny + h->seg_rows - 1) / h->seg_rows;
  }
  if (hipMemcpy(h->exn, ex.data(), sizeof(double) * ex.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->eyn, ey.data(), sizeof(double) * ey.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->exf, fx.data(), sizeof(double) * fx.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->eyf, fy.data(), sizeof(double) * fy.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->C, &C, sizeof C, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->eqa, eqa.data(), sizeof(double) * eqa.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->eqb, eqb.data(), sizeof(double) * eqb.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemsetAsync(h->ctrl, 0, sizeof(Ctrl), h->stream) != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess)
    return bail(FVWB_ERR_HIP, "copying tables failed");
  *out = h;
  return FVWB_OK;
}

static Args stage_args(fvwb_dg* h) {
  const fvwb_dg_params& p = h->p;
  Args a{};
  a.C = h->C; a.exn = h->exn; a.eyn = h->eyn; a.exf = h->exf; a.eyf = h->eyf; a.ctrl = h->ctrl;
  a.plane = h->plane; a.pitch = h->pitch; a.nx = p.nx; a.ny = p.ny;
  a.ntx = h->ntx; a.nseg = h->nseg; a.seg_rows = h->seg_rows;
  a.cfl_over = p.cfl / 5.0;
  a.gamma = (float)p.gamma; a.g = (float)p.g;
  a.odx = (float)(p.nx / p.boxlen_x); a.ody = (float)(p.ny / p.boxlen_y);
  a.rho0 = p.rho0;
  a.por = (float)(p.p0 / p.rho0);
  a.eor = (float)(p.p0 / (p.rho0 * (p.gamma - 1.0)));
  a.epor = (float)(p.p0 * p.gamma / (p.rho0 * (p.gamma - 1.0)));
  a.gpor = (float)(p.p0 * p.gamma / p.rho0);
  a.ggm1 = (float)(p.gamma * (p.gamma - 1.0));
  a.gfloor = (float)(p.gamma * 1e-10);
  a.t_end = INFINITY;
  a.dtc = 1.0;
  a.xin = nullptr; a.xout = nullptr;
  return a;
}

static LimArgs lim_args(fvwb_dg* h, float* st) {
  const fvwb_dg_params& p = h->p;
  LimArgs a{};
  a.st = st; a.exn = h->exn; a.eyn = h->eyn; a.ctrl = h->ctrl;
  a.plane = h->plane; a.pitch = h->pitch; a.nx = p.nx; a.ny = p.ny;
  a.nrec = (int)((NM * h->plane - h->org) * 4);
  a.cfl_over = p.cfl / 5.0;
  a.gamma = (float)p.gamma;
  a.odx = (float)(p.nx / p.boxlen_x); a.ody = (float)(p.ny / p.boxlen_y);
  const double dx = p.boxlen_x / p.nx, dy = p.boxlen_y / p.ny;
  a.mdx2 = (float)(p.tvb_m * (dx * dx)); a.mdy2 = (float)(p.tvb_m * (dy * dy));
  a.rho0 = p.rho0; a.eover = p.p0 / (p.gamma - 1.0);
  a.limit = p.limiter == FVWB_DG_LIMITER_TVB || p.limiter == FVWB_DG_LIMITER_TVB_POSITIVITY;
  a.charlim = p.limiter == FVWB_DG_LIMITER_CHARACTERISTIC || p.limiter == FVWB_DG_LIMITER_CHARACTERISTIC_POSITIVITY;
  a.posit = p.limiter == FVWB_DG_LIMITER_POSITIVITY || p.limiter == FVWB_DG_LIMITER_TVB_POSITIVITY ||
            p.limiter == FVWB_DG_LIMITER_CHARACTERISTIC_POSITIVITY;
  a.eqa = h->eqa; a.eqb = h->eqb;
  a.t_end = INFINITY;
  for (int q = 0; q < 3; ++q) {
    a.phi1[q] = (float)leg(h->xq[q], 1);
    a.phi2[q] = (float)leg(h->xq[q], 2);
  }
  return a;
}

static void launch_stage(fvwb_dg* h, const Args& a) {
  const dim3 grid((unsigned)((h->ntx * h->nseg + 3) / 4)), block(256);
  const bool dual = a.xout != nullptr;
  switch (h->p.flux) {
    case FVWB_FLUX_HLL:
      if (dual) hipLaunchKernelGGL((dg_stage_kernel<fvwb::FLUX_HLL, true>), grid, block, 0, h->stream, a);
      else hipLaunchKernelGGL((dg_stage_kernel<fvwb::FLUX_HLL, false>), grid, block, 0, h->stream, a);
      break;
    case FVWB_FLUX_HLLC:
      if (dual) hipLaunchKernelGGL((dg_stage_kernel<fvwb::FLUX_HLLC, true>), grid, block, 0, h->stream, a);
      else hipLaunchKernelGGL((dg_stage_kernel<fvwb::FLUX_HLLC, false>), grid, block, 0, h->stream, a);
      break;
    default:
      if (dual) hipLaunchKernelGGL((dg_stage_kernel<fvwb::FLUX_RUSANOV, true>), grid, block, 0, h->stream, a);
      else hipLaunchKernelGGL((dg_stage_kernel<fvwb::FLUX_RUSANOV, false>), grid, block, 0, h->stream, a);
  }
}

static void launch_limit(fvwb_dg* h, const LimArgs& a) {
  const dim3 grid((unsigned)((h->p.nx + 255) / 256), (unsigned)h->p.ny), block(256);
  switch (h->p.limiter) {
    case FVWB_DG_LIMITER_TVB:
      hipLaunchKernelGGL((dg_limit_kernel<true, false, false>), grid, block, 0, h->stream, a);
      break;
    case FVWB_DG_LIMITER_POSITIVITY:
      hipLaunchKernelGGL((dg_limit_kernel<false, false, true>), grid, block, 0, h->stream, a);
      break;
    case FVWB_DG_LIMITER_TVB_POSITIVITY:
      hipLaunchKernelGGL((dg_limit_kernel<true, false, true>), grid, block, 0, h->stream, a);
      break;
    case FVWB_DG_LIMITER_CHARACTERISTIC:
      hipLaunchKernelGGL((dg_limit_kernel<false, true, false>), grid, block, 0, h->stream, a);
      break;
    case FVWB_DG_LIMITER_CHARACTERISTIC_POSITIVITY:
      hipLaunchKernelGGL((dg_limit_kernel<false, true, true>), grid, block, 0, h->stream, a);
      break;
    default:
      hipLaunchKernelGGL((dg_limit_kernel<false, false, false>), grid, block, 0, h->stream, a);
  }
}

static int put_modes(fvwb_dg* h, const double* modes, float* dst) {
  const fvwb_dg_params& p = h->p;
  const long long n = 4LL * NM * p.nx * p.ny;
  DGHIP(h, hipMemcpyAsync(h->host_stage, modes, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(dg_to_planes_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, h->host_stage, dst,
                     p.nx, p.ny, h->pitch, h->plane);
  DGHIP(h, hipGetLastError());
  return FVWB_OK;
}

extern "C" int fvwb_dg_upload(fvwb_dg_handle h, const double* modes) {
  if (!h || !modes) return dg_fail(h, FVWB_ERR_ARG, "NULL argument");
  int rc = put_modes(h, modes, h->cell(0));
  if (rc) return rc;
  Ctrl z;
  std::memset(&z, 0, sizeof z);
  DGHIP(h, hipMemcpyAsync(h->ctrl, &z, sizeof z, hipMemcpyHostToDevice, h->stream));
  LimArgs la = lim_args(h, h->cell(0));
  la.parity = 0; la.cfl = 1; la.init = 1; la.bookkeeping = 0;
  launch_limit(h, la);
  DGHIP(h, hipGetLastError());
  DGHIP(h, hipStreamSynchronize(h->stream));
  h->parity = 0;
  h->ready = true;
  Ctrl c;
  DGHIP(h, hipMemcpy(&c, h->ctrl, sizeof c, hipMemcpyDeviceToHost));
  if (c.err) return dg_fail(h, FVWB_ERR_NONPHYSICAL, "initial state has rho<=0, p<=0 or non-finite values");
  return FVWB_OK;
}

extern "C" int fvwb_dg_init_fourier(fvwb_dg_handle h, int32_t nterms, const double* a, const double* mk,
                                    const double* nk, const double* theta, double amp) {
  if (!h || nterms < 0 || (nterms > 0 && (!a || !mk || !nk || !theta)))
    return dg_fail(h, FVWB_ERR_ARG, "bad Fourier arguments");
  const fvwb_dg_params& p = h->p;
  const int nx = p.nx, ny = p.ny;
  const double dx = p.boxlen_x / nx, dy = p.boxlen_y / ny, pi2 = 2.0 * std::acos(-1.0);
  double sa = 0.0;
  for (int q = 0; q < nterms; ++q) sa += std::fabs(a[q]);
  const double scale = sa > 0 ? amp / sa : 0.0;
  std::vector<double> modes(4 * NM * (size_t)nx * ny, 0.0);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i)
      for (int b = 0; b < 3; ++b)
        for (int qa = 0; qa < 3; ++qa) {
          const double x = (i + 0.5) * dx + 0.5 * dx * h->xq[qa], y = (j + 0.5) * dy + 0.5 * dy * h->xq[b];
          double dp = 0.0;
          for (int q = 0; q < nterms; ++q) dp += a[q] * std::cos(pi2 * (mk[q] * x + nk[q] * y) + theta[q]);
          const double de = dp * scale / (p.gamma - 1.0);
          const double w = 0.25 * h->wq[qa] * h->wq[b];
          for (int m = 0; m < NM; ++m)
            modes[3 + 4 * ((size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * m))] +=
                w * leg(h->xq[qa], MI[m]) * leg(h->xq[b], MJ[m]) * de;
        }
  return fvwb_dg_upload(h, modes.data());
}

extern "C" int fvwb_dg_compute_residual(fvwb_dg_handle h, const double* modes, double* dudt) {
  if (!h || !modes || !dudt) return dg_fail(h, FVWB_ERR_ARG, "NULL argument");
  int rc = put_modes(h, modes, h->cell(1));
  if (rc) return rc;
  Args a = stage_args(h);
  a.in = h->buf[1]; a.un = nullptr; a.out = h->buf[2]; a.residual_only = 1; a.parity = h->parity;
  launch_stage(h, a);
  DGHIP(h, hipGetLastError());
  const fvwb_dg_params& p = h->p;
  const long long n = 4LL * NM * p.nx * p.ny;
  hipLaunchKernelGGL(dg_from_planes_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, h->cell(2),
                     h->host_stage, p.nx, p.ny, h->pitch, h->plane);
  DGHIP(h, hipGetLastError());
  DGHIP(h, hipMemcpyAsync(dudt, h->host_stage, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
  DGHIP(h, hipStreamSynchronize(h->stream));
  return FVWB_OK;
}

// Coefficients of the SSP(5,4) of 2d/benchmark_2d_dg.f90:686-707, as the
// reference's default-real literals round them: c10, a20, a21, c21, a30, a32,
// c32, a40, a43, c43, a50, a52, a53, a54, c53, c54.
static const float SSP54[16] = {0.391752226571890f, 0.444370493651235f, 0.555629506348765f, 0.368410593050371f,
                                0.620101851488403f, 0.379898148511597f, 0.251891774271694f, 0.178079954393132f,
                                0.821920045606868f, 0.544974750228521f, 0.00683325884039f,  0.51723167208978f,
                                0.12759831133288f,  0.34833675773694f,  0.08460416338212f,  0.22600748319395f};

// Enqueue one step.  SSP-RK3: A -> B, (B,A) -> C, (C,A) -> A, each stage
// followed by the limiter; the last also reduces the CFL rate and advances the
// clock.  SSP(5,4) on four buffers: w1 = B, w2 = C, w3 = B, w4 = D; the fourth
// stage also forms w5 - c54 dt L(w4) = a50 u + a52 w2 + a53 w3 + a54 w4 + c53 dt
// L(w3) in C over w2 (each owned cell read before it is written), and the fifth
// adds a54 w4 + c54 dt L(w4) to it in A.
static int enqueue_step(fvwb_dg* h, double t_end) {
  float *A = h->buf[0], *B = h->buf[1], *Cc = h->buf[2];
  const int par = h->parity;
  Args a = stage_args(h);
  a.parity = par; a.t_end = t_end;
  LimArgs la = lim_args(h, nullptr);
  la.parity = par; la.t_end = t_end;
  if (h->p.integrator == FVWB_DG_SSPRK54) {
    const float* c = SSP54;
    float* D = h->buf[3];
    // w1 = u + c10 dt L(u)
    a.in = A; a.un = nullptr; a.out = B; a.alpha = 0.f; a.beta = 1.f; a.dtc = c[0]; a.first = 1;
    launch_stage(h, a);
    la.st = h->cell(1); la.cfl = 0; la.bookkeeping = 0;
    if (la.limit || la.charlim || la.posit) launch_limit(h, la);
    // w2 = a20 u + a21 (w1 + (c21/a21) dt L(w1))
    a.in = B; a.un = A; a.out = Cc; a.alpha = c[1]; a.beta = c[2]; a.dtc = (double)c[3] / c[2]; a.first = 0;
    launch_stage(h, a);
    la.st = h->cell(2);
    if (la.limit || la.charlim || la.posit) launch_limit(h, la);
    // w3 = a30 u + a32 (w2 + (c32/a32) dt L(w2)), over w1
    a.in = Cc; a.un = A; a.out = B; a.alpha = c[4]; a.beta = c[5]; a.dtc = (double)c[6] / c[5];
    launch_stage(h, a);
    la.st = h->cell(1);
    if (la.limit || la.charlim || la.posit) launch_limit(h, la);
    // w4 = a40 u + a43 w3 + c43 dt L(w3) and P = a50 u + a52 w2 + a53 w3 + c53 dt L(w3)
    a.in = B; a.un = A; a.out = D; a.alpha = c[7]; a.beta = c[8]; a.dtc = 1.0;
    a.xin = Cc; a.xout = Cc; a.gam = c[9]; a.pu = c[10]; a.px = c[11]; a.pin = c[12]; a.pl = c[14];
    launch_stage(h, a);
    a.xin = nullptr; a.xout = nullptr;
    la.st = h->cell(3);
    if (la.limit || la.charlim || la.posit) launch_limit(h, la);
    // u = P + a54 (w4 + (c54/a54) dt L(w4))
    a.in = D; a.un = Cc; a.out = A; a.alpha = 1.f; a.beta = c[13]; a.dtc = (double)c[15] / c[13];
    launch_stage(h, a);
    la.st = h->cell(0); la.cfl = 1; la.bookkeeping = 1;
    launch_limit(h, la);
    DGHIP(h, hipGetLastError());
    h->parity ^= 1;
    return FVWB_OK;
  }
  // stage 1
  a.in = A; a.un = nullptr; a.out = B; a.alpha = 0.f; a.beta = 1.f; a.first = 1;
  launch_stage(h, a);
  la.st = h->cell(1); la.cfl = 0; la.bookkeeping = 0;
  if (la.limit || la.charlim || la.posit) launch_limit(h, la);
  // stage 2
  a.in = B; a.un = A; a.out = Cc; a.alpha = 0.75f; a.beta = 0.25f; a.first = 0;
  launch_stage(h, a);
  la.st = h->cell(2);
  if (la.limit || la.charlim || la.posit) launch_limit(h, la);
  // stage 3 (in place over U^n: each owned cell's U^n is read by the lane that writes it, before the write)
  a.in = Cc; a.un = A; a.out = A; a.alpha = (float)(1.0 / 3.0); a.beta = (float)(2.0 / 3.0);
  launch_stage(h, a);
  la.st = h->cell(0); la.cfl = 1; la.bookkeeping = 1;
  launch_limit(h, la);
  DGHIP(h, hipGetLastError());
  h->parity ^= 1;
  return FVWB_OK;
}

extern "C" int fvwb_dg_advance_async(fvwb_dg_handle h, int32_t nsteps, double t_end) {
  if (!h) return dg_fail(h, FVWB_ERR_ARG, "NULL handle");
  if (!h->ready) return dg_fail(h, FVWB_ERR_STATE, "no state uploaded");
  if (nsteps < 0 || std::isnan(t_end)) return dg_fail(h, FVWB_ERR_ARG, "nsteps >= 0, t_end not NaN");
  for (int n = 0; n < nsteps; ++n) {
    int rc = enqueue_step(h, t_end);
    if (rc) return rc;
  }
  return FVWB_OK;
}

extern "C" int fvwb_dg_advance(fvwb_dg_handle h, int32_t nsteps, double t_end) {
  int rc = fvwb_dg_advance_async(h, nsteps, t_end);
  if (rc) return rc;
  Ctrl c;
  DGHIP(h, hipMemcpyAsync(&c, h->ctrl, sizeof c, hipMemcpyDeviceToHost, h->stream));
  DGHIP(h, hipStreamSynchronize(h->stream));
  if (c.err) return dg_fail(h, FVWB_ERR_NONPHYSICAL, "rho<=0, p<=0 or non-finite state reached");
  return FVWB_OK;
}

// ---- checkpoint / resume: the modes at t^n and the clock
extern "C" int fvwb_dg_checkpoint_save(fvwb_dg_handle h, const char* path) {
  if (!h || !path) return dg_fail(h, FVWB_ERR_ARG, "NULL argument");
  if (!h->ready) return dg_fail(h, FVWB_ERR_STATE, "no state uploaded");
  const fvwb_ckpt::Segment seg[2] = {{h->buf[0], sizeof(float) * 4 * NM * (size_t)h->plane}, {h->ctrl, sizeof(Ctrl)}};
  std::string err;
  const int rc = fvwb_ckpt::save(path, fvwb_ckpt::DG2D, &h->p, sizeof h->p, h->parity, seg, 2, h->stream, err);
  if (rc) h->err = err;
  return rc;
}

extern "C" int fvwb_dg_checkpoint_load(fvwb_dg_handle h, const char* path) {
  if (!h || !path) return dg_fail(h, FVWB_ERR_ARG, "NULL argument");
  const fvwb_ckpt::Segment seg[2] = {{h->buf[0], sizeof(float) * 4 * NM * (size_t)h->plane}, {h->ctrl, sizeof(Ctrl)}};
  std::string err;
  int32_t parity = 0;
  const int rc = fvwb_ckpt::load(path, fvwb_ckpt::DG2D, &h->p, sizeof h->p, &parity, seg, 2, h->stream, err);
  if (rc) {
    h->err = err;
    h->ready = false;
    return rc;
  }
  h->parity = parity;
  h->ready = true;
  return FVWB_OK;
}

extern "C" int fvwb_dg_download(fvwb_dg_handle h, double* modes) {
  if (!h || !modes) return dg_fail(h, FVWB_ERR_ARG, "NULL argument");
  if (!h->ready) return dg_fail(h, FVWB_ERR_STATE, "no state uploaded");
  const fvwb_dg_params& p = h->p;
  const long long n = 4LL * NM * p.nx * p.ny;
  hipLaunchKernelGGL(dg_from_planes_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, h->cell(0),
                     h->host_stage, p.nx, p.ny, h->pitch, h->plane);
  DGHIP(h, hipGetLastError());
  DGHIP(h, hipMemcpyAsync(modes, h->host_stage, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
  DGHIP(h, hipStreamSynchronize(h->stream));
  return FVWB_OK;
}

extern "C" int fvwb_dg_get_time(fvwb_dg_handle h, double* t, double* dt_last, int32_t* steps) {
  if (!h) return dg_fail(h, FVWB_ERR_ARG, "NULL handle");
  Ctrl c;
  DGHIP(h, hipMemcpyAsync(&c, h->ctrl, sizeof c, hipMemcpyDeviceToHost, h->stream));
  DGHIP(h, hipStreamSynchronize(h->stream));
  if (t) *t = c.t[h->parity];
  if (dt_last) *dt_last = c.dt_last;
  if (steps) *steps = c.steps;
  return FVWB_OK;
}

extern "C" int fvwb_dg_get_stream(fvwb_dg_handle h, void** stream) {
  if (!h || !stream) return dg_fail(h, FVWB_ERR_ARG, "NULL argument");
  *stream = (void*)h->stream;
  return FVWB_OK;
}
