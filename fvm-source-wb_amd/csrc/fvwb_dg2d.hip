// fvwb_dg2d.hip — 2D modal DG (P2, 3x3 Gauss-Legendre) in deviation form on MI355X.
//
// Residual of 2d/benchmark_2d_dg.f90:1137-1479 (volume integrals :1207-1244,
// traces :1253-1314, face fluxes :1333-1366, edge integrals :1372-1411, source
// :1426-1443, assembly :1455-1462) written for the deviation dU = U - U_eq of
// dg_with_source.f90:1749-2031: all fluxes and the source are deviations from
// the equilibrium's, so a resting atmosphere is an exact fixed point.
//
// The projections are small dense products per cell and run on the matrix
// cores as v_mfma_f32_16x16x4f32 with 16 cells in the N dimension:
//   forward:  [16 points x 8 modes] x [8 modes x 16 cells]  (volume nodes, and
//             separately the 12 face points)
//   backward: [16 modes x 48 inputs] x [48 inputs x 16 cells] for the 27 volume
//             inputs (F_x/dx, F_y/dy, S at 9 nodes, zero-padded to 16 each) and
//             [16 x 12] x [12 x 16] for the face fluxes.
// Lane l of the output holds points (or modes) 4*(l/16)+r of cell l%16, so the
// point-wise flux work is lane-local and each lane feeds its own outputs into
// the backward product as its k-slot (the k order of a product is free).
// A workgroup owns a 16x16 tile; its 1-cell halo only contributes face traces.
// Face traces are exchanged through LDS, where the face fluxes are computed.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/fvwb.h"
#include "fvwb_device.h"

namespace fvwb {
namespace dg {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int NM = 6;      // P2 modes
constexpr int NQ = 3;      // Gauss points per direction
constexpr int TILE = 16;   // cells per tile side

// Per-lane MFMA operand constants, built on the host from the basis.
struct Consts {
  float fwd_node[16][8];   // V[row][mode], row 4g + r -> Gauss node g + 4r (node a + 3b; 9 used)
  float fwd_face[16][8];   // trace points: 0-2 left (xi=-1, eta_b), 3-5 right, 6-8 bottom (xi_a, eta=-1), 9-11 top
  float back_vol[3][16][16];  // [quantity][mode][row as fwd_node]: 0.5 w dphi_i phi_j (F_x), 0.5 w phi_i dphi_j (F_y), 0.25 w phi phi (S)
  float back_face[16][12];    // [mode][face point]: +-0.5 w phi (flux pre-scaled by 1/dx or 1/dy)
};

struct Ctrl {
  uint32_t rate[2];   // FP32 bit patterns of max((|u|+c)/dx + (|v|+c)/dy); positive floats order like uints
  double t[2];
  double dt_last;
  int32_t steps;
  uint32_t err;       // bit 0: rho<=0, p<=0 or non-finite seen
};

struct Args {
  const float* in;     // stage input planes (24 = 4 vars x 6 modes)
  const float* un;     // U^n planes (stage combination), may be null when alpha == 0
  float* out;          // output planes (may alias un)
  const Consts* C;
  const double* exn;   // e^{-k x} at x nodes   [3*nx]
  const double* eyn;   // e^{-k y} at y nodes   [3*ny]
  const double* exf;   // e^{-k x} at x faces   [nx+1]
  const double* eyf;   // e^{-k y} at y faces   [ny+1]
  Ctrl* ctrl;
  long long plane;
  int pitch, nx, ny, parity;
  int first, residual_only;
  double t_end, cfl_over;   // cfl / (2k+1)
  float alpha, beta;
  float gamma, g, odx, ody;
  double rho0, p0, gm1d;
};

__device__ inline float ld_mode(const float* base, long long plane, int pitch, int v, int m, int i, int j) {
  return base[(long long)(v * NM + m) * plane + (long long)j * pitch + i];
}

// Deviation fluxes along x and y and the wave-speed rate at one point.
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

// Equilibrium at a point from the separable FP64 tables, rounded once to FP32.
__device__ inline FaceEq<float> eq_at(const Args& a, double ex, double ey) {
  const double e = ex * ey, pe = a.p0 * e, ee = pe / a.gm1d;
  FaceEq<float> q;
  q.rho = (float)(a.rho0 * e);
  q.p = (float)pe;
  q.e = (float)ee;
  q.ep = (float)(ee + pe);
  return q;
}

// Lane-local node assignment of the forward volume product: row 4*g + r of the
// node block (g = lane/16) is Gauss node g + 4r, so the nine nodes spread as
// 3,2,2,2 over the four lane groups and the backward volume product needs only
// r = 0..2 (9 steps of 4 k-slots) instead of 12.
__host__ __device__ constexpr int node_of_row(int row) { return (row >> 2) + 4 * (row & 3); }

template <int FLUX>
__global__ __launch_bounds__(256) void dg_stage_kernel(Args a) {
  const Ctrl* ck = a.ctrl;
  const int par = a.parity;
  const double t = ck->t[par];
  if (!a.residual_only && !(t < a.t_end)) return;       // this instance is done
  double dt = 0.0;
  if (!a.residual_only) {
    const float rate = __uint_as_float(ck->rate[par]);
    dt = a.cfl_over / (double)rate;
    if (t + dt > a.t_end) dt = a.t_end - t;
    if (a.first && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) a.ctrl->rate[par ^ 1] = 0u;
  }
  const float dtf = (float)dt;

  // face traces, then face fluxes: the innermost index runs along the lanes
  __shared__ float XT[2][4][TILE][NQ][TILE + 1];   // x faces: [side][var][row][point][face]
  __shared__ float YT[2][4][TILE + 1][NQ][TILE];   // y faces: [side][var][face][point][col]
  // equilibrium tables of the tile (halo included): e^{-kx}, e^{-ky} at nodes and faces
  __shared__ double Exn[3 * TILE], Eyn[3 * TILE], Exf[TILE + 1], Eyf[TILE + 1];

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c16 = lane & 15, g4 = lane >> 4;
  const int tx0 = blockIdx.x * TILE, ty0 = blockIdx.y * TILE;
  const Consts& C = *a.C;

  {
    const int q = threadIdx.x;
    if (q < 3 * TILE) {
      const int ci = min(tx0 + q / 3, a.nx - 1), cj = min(ty0 + q / 3, a.ny - 1);
      Exn[q] = a.exn[3 * ci + q % 3];
      Eyn[q] = a.eyn[3 * cj + q % 3];
    } else if (q < 3 * TILE + TILE + 1) {
      const int f = q - 3 * TILE;
      Exf[f] = a.exf[min(tx0 + f, a.nx)];
      Eyf[f] = a.eyf[min(ty0 + f, a.ny)];
    }
  }

  Phys<float> P;
  P.gamma = a.gamma; P.gm1 = a.gamma - 1.0f; P.half = 0.5f; P.g = a.g;

  // all mode loads of the tile up front: lane supplies mode g4 (+4) of cell c16
  float B0[4][4], B1[4][4], N0[4][4], N1[4][4];   // [row group][var]: stage input and U^n
  const int i = tx0 + c16;
#pragma unroll
  for (int rr = 0; rr < 4; ++rr) {
    const int j = ty0 + 4 * wave + rr;
    const bool ok = j < a.ny && i < a.nx;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      B0[rr][v] = ok ? ld_mode(a.in, a.plane, a.pitch, v, g4, i, j) : 0.f;
      B1[rr][v] = (ok && g4 + 4 < NM) ? ld_mode(a.in, a.plane, a.pitch, v, g4 + 4, i, j) : 0.f;
      const bool okn = ok && a.alpha != 0.f;
      N0[rr][v] = okn ? ld_mode(a.un, a.plane, a.pitch, v, g4, i, j) : 0.f;
      N1[rr][v] = (okn && g4 + 4 < NM) ? ld_mode(a.un, a.plane, a.pitch, v, g4 + 4, i, j) : 0.f;
    }
  }
  // halo cell of this wave: below, above, left or right of the tile
  int hi, hj;
  bool hok;
  if (wave == 0) { hi = tx0 + c16; hj = ty0 - 1; hok = ty0 >= 1; }
  else if (wave == 1) { hi = tx0 + c16; hj = ty0 + TILE; hok = ty0 + TILE < a.ny; }
  else if (wave == 2) { hi = tx0 - 1; hj = ty0 + c16; hok = tx0 >= 1; }
  else { hi = tx0 + TILE; hj = ty0 + c16; hok = tx0 + TILE < a.nx; }
  hok = hok && hi < a.nx && hj < a.ny;
  float H0[4], H1[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    H0[v] = hok ? ld_mode(a.in, a.plane, a.pitch, v, g4, hi, hj) : 0.f;
    H1[v] = (hok && g4 + 4 < NM) ? ld_mode(a.in, a.plane, a.pitch, v, g4 + 4, hi, hj) : 0.f;
  }

  // forward operands: lane supplies V[row = lane%16][mode = lane/16 + 4s]
  const float An0 = C.fwd_node[c16][g4], An1 = C.fwd_node[c16][g4 + 4];
  const float Af0 = C.fwd_face[c16][g4], Af1 = C.fwd_face[c16][g4 + 4];
  // backward operands: lane supplies B[mode = lane%16][row = 4*(lane/16) + r]
  float Abv[3][3], Abf[3];
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int r = 0; r < 3; ++r) Abv[q][r] = C.back_vol[q][c16][4 * g4 + r];
#pragma unroll
  for (int s = 0; s < 3; ++s) Abf[s] = C.back_face[c16][4 * s + g4];

  __syncthreads();   // equilibrium tables

  v4f acc[4][4];   // [row group][var]: D[mode 4*(l/16)+r][cell l%16]

  // ---- phase 1: interior rows (volume terms + traces) --------------------------------
#pragma unroll
  for (int rr = 0; rr < 4; ++rr) {
    const int jl = 4 * wave + rr;
    v4f dn[4], df[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      v4f z = {0.f, 0.f, 0.f, 0.f};
      dn[v] = __builtin_amdgcn_mfma_f32_16x16x4f32(An0, B0[rr][v], z, 0, 0, 0);
      dn[v] = __builtin_amdgcn_mfma_f32_16x16x4f32(An1, B1[rr][v], dn[v], 0, 0, 0);
      df[v] = __builtin_amdgcn_mfma_f32_16x16x4f32(Af0, B0[rr][v], z, 0, 0, 0);
      df[v] = __builtin_amdgcn_mfma_f32_16x16x4f32(Af1, B1[rr][v], df[v], 0, 0, 0);
    }
    // traces to LDS: lane holds trace points 4*g4 + r of cell c16
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int pt = 4 * g4 + r;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float val = df[v][r];
        if (pt < 3) XT[1][v][jl][pt][c16] = val;                     // left trace -> x face c16, right side
        else if (pt < 6) XT[0][v][jl][pt - 3][c16 + 1] = val;        // right trace -> x face c16+1, left side
        else if (pt < 9) YT[1][v][jl][pt - 6][c16] = val;            // bottom trace -> y face jl, upper side
        else if (pt < 12) YT[0][v][jl + 1][pt - 9][c16] = val;       // top trace -> y face jl+1, lower side
      }
    }
    // volume terms at the lane's nodes g4, g4+4, g4+8
    float X[3][3][4];   // [quantity][r][var]
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const int node = g4 + 4 * r;
      if (node < 9) {
        const int qa = node % 3, qb = node / 3;
        const double e = Exn[3 * c16 + qa] * Eyn[3 * jl + qb], pe = a.p0 * e, ee = pe / a.gm1d;
        FaceEq<float> q;
        q.rho = (float)(a.rho0 * e); q.p = (float)pe; q.e = (float)ee; q.ep = (float)(ee + pe);
        Vec4<float> d = {dn[0][r], dn[1][r], dn[2][r], dn[3][r]};
        Vec4<float> fx, fy;
        point_terms(P, q, d, fx, fy);
        X[0][r][0] = fx.r * a.odx; X[0][r][1] = fx.mx * a.odx; X[0][r][2] = fx.my * a.odx; X[0][r][3] = fx.e * a.odx;
        X[1][r][0] = fy.r * a.ody; X[1][r][1] = fy.mx * a.ody; X[1][r][2] = fy.my * a.ody; X[1][r][3] = fy.e * a.ody;
        X[2][r][0] = 0.f; X[2][r][1] = -a.g * d.r; X[2][r][2] = -a.g * d.r; X[2][r][3] = -a.g * (d.mx + d.my);
      } else {
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
          for (int v = 0; v < 4; ++v) X[q][r][v] = 0.f;
      }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      v4f z = {0.f, 0.f, 0.f, 0.f};
      acc[rr][v] = z;
#pragma unroll
      for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int r = 0; r < 3; ++r)
          acc[rr][v] = __builtin_amdgcn_mfma_f32_16x16x4f32(Abv[q][r], X[q][r][v], acc[rr][v], 0, 0, 0);
    }
  }

  // ---- phase 1b: halo traces ------------------------------------------------------
  {
    v4f df[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      v4f z = {0.f, 0.f, 0.f, 0.f};
      df[v] = __builtin_amdgcn_mfma_f32_16x16x4f32(Af0, H0[v], z, 0, 0, 0);
      df[v] = __builtin_amdgcn_mfma_f32_16x16x4f32(Af1, H1[v], df[v], 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int pt = 4 * g4 + r;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float val = df[v][r];
        if (wave == 0 && pt >= 9 && pt < 12) YT[0][v][0][pt - 9][c16] = val;       // top trace of the row below
        if (wave == 1 && pt >= 6 && pt < 9) YT[1][v][TILE][pt - 6][c16] = val;     // bottom trace of the row above
        if (wave == 2 && pt >= 3 && pt < 6) XT[0][v][c16][pt - 3][0] = val;        // right trace of the left column
        if (wave == 3 && pt < 3) XT[1][v][c16][pt][TILE] = val;                     // left trace of the right column
      }
    }
  }
  __syncthreads();

  // ---- phase 2: face fluxes (minus equilibrium flux), scaled by 1/dx or 1/dy ----------
  for (int it = threadIdx.x; it < (TILE + 1) * TILE * NQ; it += 256) {
    // x face f (innermost), row, point
    {
      const int f = it % (TILE + 1), pt = (it / (TILE + 1)) % NQ, row = it / ((TILE + 1) * NQ);
      const int fi = tx0 + f, j = ty0 + row;
      if (fi <= a.nx && j < a.ny) {
        const bool hasl = fi >= 1, hasr = fi <= a.nx - 1;
        Vec4<float> dl = {XT[0][0][row][pt][f], XT[0][1][row][pt][f], XT[0][2][row][pt][f], XT[0][3][row][pt][f]};
        Vec4<float> dr = {XT[1][0][row][pt][f], XT[1][1][row][pt][f], XT[1][2][row][pt][f], XT[1][3][row][pt][f]};
        if (!hasl) dl = {dr.r, -dr.mx, dr.my, dr.e};
        if (!hasr) dr = {dl.r, -dl.mx, dl.my, dl.e};
        const double e = Exf[f] * Eyn[3 * row + pt], pe = a.p0 * e, ee = pe / a.gm1d;
        FaceEq<float> q;
        q.rho = (float)(a.rho0 * e); q.p = (float)pe; q.e = (float)ee; q.ep = (float)(ee + pe);
        Vec4<float> F = face_flux<float, FLUX, 0, FORM_DEV>(P, q, dl, dr);
        XT[0][0][row][pt][f] = F.r * a.odx; XT[0][1][row][pt][f] = F.mx * a.odx;
        XT[0][2][row][pt][f] = F.my * a.odx; XT[0][3][row][pt][f] = F.e * a.odx;
      }
    }
    // y face f, point, column (innermost)
    {
      const int col = it % TILE, pt = (it / TILE) % NQ, f = it / (TILE * NQ);
      const int fj = ty0 + f, ii = tx0 + col;
      if (fj <= a.ny && ii < a.nx) {
        const bool hasb = fj >= 1, hast = fj <= a.ny - 1;
        Vec4<float> dl = {YT[0][0][f][pt][col], YT[0][1][f][pt][col], YT[0][2][f][pt][col], YT[0][3][f][pt][col]};
        Vec4<float> dr = {YT[1][0][f][pt][col], YT[1][1][f][pt][col], YT[1][2][f][pt][col], YT[1][3][f][pt][col]};
        if (!hasb) dl = {dr.r, dr.mx, -dr.my, dr.e};
        if (!hast) dr = {dl.r, dl.mx, -dl.my, dl.e};
        const double e = Exn[3 * col + pt] * Eyf[f], pe = a.p0 * e, ee = pe / a.gm1d;
        FaceEq<float> q;
        q.rho = (float)(a.rho0 * e); q.p = (float)pe; q.e = (float)ee; q.ep = (float)(ee + pe);
        Vec4<float> F = face_flux<float, FLUX, 1, FORM_DEV>(P, q, dl, dr);
        YT[0][0][f][pt][col] = F.r * a.ody; YT[0][1][f][pt][col] = F.mx * a.ody;
        YT[0][2][f][pt][col] = F.my * a.ody; YT[0][3][f][pt][col] = F.e * a.ody;
      }
    }
  }
  __syncthreads();

  // ---- phase 3: face integrals, stage combination, store -----------------------------
#pragma unroll
  for (int rr = 0; rr < 4; ++rr) {
    const int jl = 4 * wave + rr, j = ty0 + jl;
    const bool ok = j < a.ny && i < a.nx;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        const int k = 4 * s + g4;   // face point fed by this lane in step s
        float b;
        if (k < 3) b = XT[0][v][jl][k][c16];
        else if (k < 6) b = XT[0][v][jl][k - 3][c16 + 1];
        else if (k < 9) b = YT[0][v][jl][k - 6][c16];
        else b = YT[0][v][jl + 1][k - 9][c16];
        if (!ok) b = 0.f;
        acc[rr][v] = __builtin_amdgcn_mfma_f32_16x16x4f32(Abf[s], b, acc[rr][v], 0, 0, 0);
      }
    }
    // U and U^n in the output layout: mode 4*g4 + r of cell c16 sits in lane
    // 16*r + c16, as its first (g4 == 0) or second (g4 == 1) loaded mode
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int src = 16 * r + c16;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float u0 = __shfl(B0[rr][v], src), u1 = __shfl(B1[rr][v], src);
        const float n0 = __shfl(N0[rr][v], src), n1 = __shfl(N1[rr][v], src);
        const int m = 4 * g4 + r;
        if (ok && g4 < 2 && m < NM) {
          const float u = g4 == 0 ? u0 : u1, un = g4 == 0 ? n0 : n1;
          const float L = acc[rr][v][r];
          float res;
          if (a.residual_only) {
            res = L;
          } else {
            res = a.beta * (u + dtf * L);
            if (a.alpha != 0.f) res = a.alpha * un + res;
          }
          a.out[(long long)(v * NM + m) * a.plane + (long long)j * a.pitch + i] = res;
        }
      }
    }
  }
}

// TVB limiter (in place; means are never written, so neighbours may be read
// concurrently) and, when requested, the CFL rate of the limited state plus the
// step bookkeeping of the instance.
struct LimArgs {
  float* st;
  const double* exn;
  const double* eyn;
  Ctrl* ctrl;
  long long plane;
  int pitch, nx, ny, parity;
  int limit, cfl, bookkeeping, init;
  double t_end, cfl_over;
  float gamma, odx, ody, mdx2, mdy2;
  double rho0, p0;
  float phi[3][3];   // phi_n(xi_a)
};

__device__ inline float gminmod(float x, float y, float z, float mh2) {
  if (fabsf(x) < mh2) return x;
  const bool sx = signbit(x);
  if (signbit(y) != sx || signbit(z) != sx) return 0.f;
  const float m = fminf(fminf(fabsf(x), fabsf(y)), fabsf(z));
  return sx ? -m : m;
}

__global__ __launch_bounds__(256) void dg_limit_kernel(LimArgs a) {
  const int par = a.parity;
  const bool active = a.init || (a.ctrl->t[par] < a.t_end);
  if (!active) {
    // instance done: carry the clock and rate to the next parity once
    if (a.bookkeeping && blockIdx.x == 0 && threadIdx.x == 0) {
      a.ctrl->t[par ^ 1] = a.ctrl->t[par];
      a.ctrl->rate[par ^ 1] = a.ctrl->rate[par];
    }
    return;
  }
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int nc = a.nx * a.ny;
  float rate = 0.f;
  bool bad = false;
  if (c < nc) {
    const int i = c % a.nx, j = c / a.nx;
    const long long o = (long long)j * a.pitch + i;
    float md[4][NM];
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int m = 0; m < NM; ++m) md[v][m] = a.st[(long long)(v * NM + m) * a.plane + o];
    if (a.limit) {
      const float s3 = 1.7320508075688772f;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float* base = a.st + (long long)(v * NM) * a.plane;
        const float u0 = md[v][0];
        const float uw = i > 0 ? base[o - 1] : (v == 1 ? -u0 : u0);
        const float ue = i < a.nx - 1 ? base[o + 1] : (v == 1 ? -u0 : u0);
        const float us = j > 0 ? base[o - a.pitch] : (v == 2 ? -u0 : u0);
        const float un = j < a.ny - 1 ? base[o + a.pitch] : (v == 2 ? -u0 : u0);
        const float x = s3 * md[v][1], y = s3 * md[v][2];
        const float xt = gminmod(x, ue - u0, u0 - uw, a.mdx2);
        const float yt = gminmod(y, un - u0, u0 - us, a.mdy2);
        if (xt != x || yt != y) {
          md[v][1] = xt * (1.0f / 1.7320508075688772f);
          md[v][2] = yt * (1.0f / 1.7320508075688772f);
          md[v][3] = md[v][4] = md[v][5] = 0.f;
          for (int m = 1; m < NM; ++m) a.st[(long long)(v * NM + m) * a.plane + o] = md[v][m];
        }
      }
    }
    if (a.cfl) {
      const float gm1 = a.gamma - 1.0f;
      const double e_over = a.p0 / (double)gm1;
      double ex[NQ], ey[NQ];
#pragma unroll
      for (int q = 0; q < NQ; ++q) { ex[q] = a.exn[3 * i + q]; ey[q] = a.eyn[3 * j + q]; }
#pragma unroll
      for (int qb = 0; qb < NQ; ++qb)
#pragma unroll
        for (int qa = 0; qa < NQ; ++qa) {
          const double e = ex[qa] * ey[qb];
          const float re = (float)(a.rho0 * e), ee = (float)(e_over * e);
          float d[4];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const float* mm = md[v];
            d[v] = mm[0] + a.phi[1][qa] * mm[1] + a.phi[1][qb] * mm[2] + a.phi[2][qa] * mm[3] +
                   a.phi[1][qa] * a.phi[1][qb] * mm[4] + a.phi[2][qb] * mm[5];
          }
          const float rho = re + d[0];
          const float inv = trcp(fmaxf(rho, 1e-10f));
          const float u = d[1] * inv, w = d[2] * inv;
          const float p = gm1 * ((ee + d[3]) - 0.5f * rho * (u * u + w * w));
          const float cs = tsqrt(a.gamma * fmaxf(p, 1e-10f) * inv);
          const float rt = (fabsf(u) + cs) * a.odx + (fabsf(w) + cs) * a.ody;
          bad |= !(rho > 0.f) || !(p > 0.f) || !(rt < 1e30f);
          rate = fmaxf(rate, rt);
        }
    }
  }
  if (a.cfl) {
    for (int off = 32; off > 0; off >>= 1) rate = fmaxf(rate, __shfl_xor(rate, off));
    const int anybad = __any(bad);
    const int slot = a.init ? par : (par ^ 1);
    if ((threadIdx.x & 63) == 0) {
      atomicMax(&a.ctrl->rate[slot], __float_as_uint(rate));
      if (anybad) atomicOr(&a.ctrl->err, 1u);
    }
    if (a.bookkeeping && blockIdx.x == 0 && threadIdx.x == 0) {
      // the step that just ended: same dt as its stages computed
      const double t = a.ctrl->t[par];
      double dt = a.cfl_over / (double)__uint_as_float(a.ctrl->rate[par]);
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
  float* buf[3] = {nullptr, nullptr, nullptr};
  double* host_stage = nullptr;   // device FP64 AoS staging
  double *exn = nullptr, *eyn = nullptr, *exf = nullptr, *eyf = nullptr;
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
  void* ps[] = {h->buf[0], h->buf[1], h->buf[2], h->host_stage, h->exn, h->eyn, h->exf, h->eyf, h->C, h->ctrl};
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
  if (p->flux != FVWB_FLUX_RUSANOV && p->flux != FVWB_FLUX_HLL) return dg_fail(nullptr, FVWB_ERR_ARG, "bad flux");
  if (p->limiter != FVWB_DG_LIMITER_NONE && p->limiter != FVWB_DG_LIMITER_TVB)
    return dg_fail(nullptr, FVWB_ERR_ARG, "bad limiter");
  if (p->bc != FVWB_BC_REFLECT) return dg_fail(nullptr, FVWB_ERR_ARG, "the DG path implements walls (FVWB_BC_REFLECT)");
  if (!(p->gamma > 1.0) || !(p->boxlen_x > 0) || !(p->boxlen_y > 0) || !(p->cfl > 0) || !(p->rho0 > 0) ||
      !(p->p0 > 0) || !(p->g >= 0) || !(p->tvb_m >= 0))
    return dg_fail(nullptr, FVWB_ERR_ARG, "gamma > 1, positive box, cfl, rho0, p0 and g, M >= 0 required");
  if ((long long)p->nx * p->ny > (1LL << 26)) return dg_fail(nullptr, FVWB_ERR_ARG, "nx*ny must be <= 2^26");
  fvwb_dg* h = new (std::nothrow) fvwb_dg();
  if (!h) return dg_fail(nullptr, FVWB_ERR_OOM, "host allocation failed");
  h->p = *p;
  const int nx = p->nx, ny = p->ny;
  h->pitch = (nx + TILE - 1) / TILE * TILE;
  h->plane = (long long)h->pitch * ((ny + TILE - 1) / TILE * TILE);
  auto bail = [&](int rc, const char* m) { h->err = m; g_dg_err = m; fvwb_dg_destroy(h); return rc; };
  if (hipStreamCreate(&h->stream) != hipSuccess) return bail(FVWB_ERR_HIP, "hipStreamCreate failed");
  const size_t sbytes = sizeof(float) * 4 * NM * (size_t)h->plane;
  for (int b = 0; b < 3; ++b)
    if (hipMalloc(&h->buf[b], sbytes) != hipSuccess) return bail(FVWB_ERR_OOM, "device allocation failed");
  if (hipMalloc(&h->host_stage, sizeof(double) * 4 * NM * (size_t)nx * ny) != hipSuccess ||
      hipMalloc(&h->exn, sizeof(double) * 3 * nx) != hipSuccess || hipMalloc(&h->eyn, sizeof(double) * 3 * ny) != hipSuccess ||
      hipMalloc(&h->exf, sizeof(double) * (nx + 1)) != hipSuccess ||
      hipMalloc(&h->eyf, sizeof(double) * (ny + 1)) != hipSuccess || hipMalloc(&h->C, sizeof(Consts)) != hipSuccess ||
      hipMalloc(&h->ctrl, sizeof(Ctrl)) != hipSuccess)
    return bail(FVWB_ERR_OOM, "device allocation failed");
  // padding columns and rows stay zero for the whole life of the handle
  for (int b = 0; b < 3; ++b) {
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
  Consts C;
  std::memset(&C, 0, sizeof C);
  for (int m = 0; m < NM; ++m) {
    for (int row = 0; row < 16; ++row) {
      const int node = node_of_row(row);
      if (node >= 9) continue;
      const int a = node % 3, b = node / 3;
      const double w = h->wq[a] * h->wq[b];
      const double pa = leg(h->xq[a], MI[m]), pb = leg(h->xq[b], MJ[m]);
      C.fwd_node[row][m] = (float)(pa * pb);
      C.back_vol[0][m][row] = (float)(0.5 * w * legp(h->xq[a], MI[m]) * pb);
      C.back_vol[1][m][row] = (float)(0.5 * w * pa * legp(h->xq[b], MJ[m]));
      C.back_vol[2][m][row] = (float)(0.25 * w * pa * pb);
    }
    for (int q = 0; q < 3; ++q) {
      C.fwd_face[q][m] = (float)(leg(-1.0, MI[m]) * leg(h->xq[q], MJ[m]));       // left
      C.fwd_face[3 + q][m] = (float)(leg(1.0, MI[m]) * leg(h->xq[q], MJ[m]));    // right
      C.fwd_face[6 + q][m] = (float)(leg(h->xq[q], MI[m]) * leg(-1.0, MJ[m]));   // bottom
      C.fwd_face[9 + q][m] = (float)(leg(h->xq[q], MI[m]) * leg(1.0, MJ[m]));    // top
      C.back_face[m][q] = (float)(0.5 * h->wq[q] * leg(-1.0, MI[m]) * leg(h->xq[q], MJ[m]));
      C.back_face[m][3 + q] = (float)(-0.5 * h->wq[q] * leg(1.0, MI[m]) * leg(h->xq[q], MJ[m]));
      C.back_face[m][6 + q] = (float)(0.5 * h->wq[q] * leg(h->xq[q], MI[m]) * leg(-1.0, MJ[m]));
      C.back_face[m][9 + q] = (float)(-0.5 * h->wq[q] * leg(h->xq[q], MI[m]) * leg(1.0, MJ[m]));
    }
  }
  if (hipMemcpy(h->exn, ex.data(), sizeof(double) * ex.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->eyn, ey.data(), sizeof(double) * ey.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->exf, fx.data(), sizeof(double) * fx.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->eyf, fy.data(), sizeof(double) * fy.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->C, &C, sizeof C, hipMemcpyHostToDevice) != hipSuccess ||
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
  a.cfl_over = p.cfl / 5.0;
  a.gamma = (float)p.gamma; a.g = (float)p.g;
  a.odx = (float)(p.nx / p.boxlen_x); a.ody = (float)(p.ny / p.boxlen_y);
  a.rho0 = p.rho0; a.p0 = p.p0; a.gm1d = p.gamma - 1.0;
  a.t_end = INFINITY;
  return a;
}

static LimArgs lim_args(fvwb_dg* h, float* st) {
  const fvwb_dg_params& p = h->p;
  LimArgs a{};
  a.st = st; a.exn = h->exn; a.eyn = h->eyn; a.ctrl = h->ctrl;
  a.plane = h->plane; a.pitch = h->pitch; a.nx = p.nx; a.ny = p.ny;
  a.cfl_over = p.cfl / 5.0;
  a.gamma = (float)p.gamma;
  a.odx = (float)(p.nx / p.boxlen_x); a.ody = (float)(p.ny / p.boxlen_y);
  const double dx = p.boxlen_x / p.nx, dy = p.boxlen_y / p.ny;
  a.mdx2 = (float)(p.tvb_m * (dx * dx)); a.mdy2 = (float)(p.tvb_m * (dy * dy));
  a.rho0 = p.rho0; a.p0 = p.p0;
  a.limit = p.limiter == FVWB_DG_LIMITER_TVB;
  a.t_end = INFINITY;
  for (int n = 0; n < 3; ++n)
    for (int q = 0; q < 3; ++q) a.phi[n][q] = (float)leg(h->xq[q], n);
  return a;
}

static void launch_stage(fvwb_dg* h, const Args& a) {
  const dim3 grid((h->p.nx + TILE - 1) / TILE, (h->p.ny + TILE - 1) / TILE), block(256);
  if (h->p.flux == FVWB_FLUX_HLL)
    hipLaunchKernelGGL(dg_stage_kernel<fvwb::FLUX_HLL>, grid, block, 0, h->stream, a);
  else
    hipLaunchKernelGGL(dg_stage_kernel<fvwb::FLUX_RUSANOV>, grid, block, 0, h->stream, a);
}

static void launch_limit(fvwb_dg* h, const LimArgs& a) {
  const int nc = h->p.nx * h->p.ny;
  hipLaunchKernelGGL(dg_limit_kernel, dim3((nc + 255) / 256), dim3(256), 0, h->stream, a);
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
  int rc = put_modes(h, modes, h->buf[0]);
  if (rc) return rc;
  Ctrl z;
  std::memset(&z, 0, sizeof z);
  DGHIP(h, hipMemcpyAsync(h->ctrl, &z, sizeof z, hipMemcpyHostToDevice, h->stream));
  LimArgs la = lim_args(h, h->buf[0]);
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
  int rc = put_modes(h, modes, h->buf[1]);
  if (rc) return rc;
  Args a = stage_args(h);
  a.in = h->buf[1]; a.un = nullptr; a.out = h->buf[2]; a.residual_only = 1; a.parity = h->parity;
  launch_stage(h, a);
  DGHIP(h, hipGetLastError());
  const fvwb_dg_params& p = h->p;
  const long long n = 4LL * NM * p.nx * p.ny;
  hipLaunchKernelGGL(dg_from_planes_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, h->buf[2],
                     h->host_stage, p.nx, p.ny, h->pitch, h->plane);
  DGHIP(h, hipGetLastError());
  DGHIP(h, hipMemcpyAsync(dudt, h->host_stage, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
  DGHIP(h, hipStreamSynchronize(h->stream));
  return FVWB_OK;
}

// Enqueue one SSP-RK3 step: A -> B, (B,A) -> C, (C,A) -> A, each stage followed
// by the limiter; the last also reduces the CFL rate and advances the clock.
static int enqueue_step(fvwb_dg* h, double t_end) {
  float *A = h->buf[0], *B = h->buf[1], *Cc = h->buf[2];
  const int par = h->parity;
  Args a = stage_args(h);
  a.parity = par; a.t_end = t_end;
  LimArgs la = lim_args(h, nullptr);
  la.parity = par; la.t_end = t_end;
  // stage 1
  a.in = A; a.un = nullptr; a.out = B; a.alpha = 0.f; a.beta = 1.f; a.first = 1;
  launch_stage(h, a);
  la.st = B; la.cfl = 0; la.bookkeeping = 0;
  if (la.limit) launch_limit(h, la);
  // stage 2
  a.in = B; a.un = A; a.out = Cc; a.alpha = 0.75f; a.beta = 0.25f; a.first = 0;
  launch_stage(h, a);
  la.st = Cc;
  if (la.limit) launch_limit(h, la);
  // stage 3 (in place over U^n: each lane reads its own U^n entry before writing it)
  a.in = Cc; a.un = A; a.out = A; a.alpha = (float)(1.0 / 3.0); a.beta = (float)(2.0 / 3.0);
  launch_stage(h, a);
  la.st = A; la.cfl = 1; la.bookkeeping = 1;
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

extern "C" int fvwb_dg_download(fvwb_dg_handle h, double* modes) {
  if (!h || !modes) return dg_fail(h, FVWB_ERR_ARG, "NULL argument");
  if (!h->ready) return dg_fail(h, FVWB_ERR_STATE, "no state uploaded");
  const fvwb_dg_params& p = h->p;
  const long long n = 4LL * NM * p.nx * p.ny;
  hipLaunchKernelGGL(dg_from_planes_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, h->buf[0],
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
