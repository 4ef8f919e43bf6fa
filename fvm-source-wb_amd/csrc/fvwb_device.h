// fvwb_device.h — device-side building blocks shared by the FV kernels.
//
// Pointwise physics of the Euler-with-gravity residual in deviation form
// (benchmark_2d.f90:145-171 primitive/conservative, 283-295 speed, 299-325 flux,
// 327-350 source, 353-367 Rusanov) and the cross-lane moves used by the
// sliding-window stage kernel.  Two flux forms:
//   * PLAIN (FP64 drop-in): full states U_eq(face)+dU, full fluxes, minus the
//     equilibrium flux — the operation structure of compute_update_exact;
//   * DEV (FP32 bulk): the flux deviation written algebraically for an
//     equilibrium at rest, so no O(1) quantity is ever subtracted from another.
// Both give exactly zero at dU == 0 under FMA contraction (see comments).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fvwb {

// ---- cross-lane moves (wave64) -------------------------------------------------
// lane l receives the value of lane l-1; lane 0 receives 0 (bound_ctrl), which
// lets the compiler fold the move into the consuming VALU op as a DPP operand.
// Lanes at the wave edges are halo lanes whose results are never stored.
__device__ __forceinline__ float lane_from_left(float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(
      0, __builtin_bit_cast(int, x), 0x138 /*wave_shr:1*/, 0xF, 0xF, true));
}
// lane l receives the value of lane l+1; lane 63 receives 0.
__device__ __forceinline__ float lane_from_right(float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(
      0, __builtin_bit_cast(int, x), 0x130 /*wave_shl:1*/, 0xF, 0xF, true));
}
__device__ __forceinline__ double lane_from_left(double x) {
  int2 v = __builtin_bit_cast(int2, x);
  v.x = __builtin_amdgcn_update_dpp(0, v.x, 0x138, 0xF, 0xF, true);
  v.y = __builtin_amdgcn_update_dpp(0, v.y, 0x138, 0xF, 0xF, true);
  return __builtin_bit_cast(double, v);
}
__device__ __forceinline__ double lane_from_right(double x) {
  int2 v = __builtin_bit_cast(int2, x);
  v.x = __builtin_amdgcn_update_dpp(0, v.x, 0x130, 0xF, 0xF, true);
  v.y = __builtin_amdgcn_update_dpp(0, v.y, 0x130, 0xF, 0xF, true);
  return __builtin_bit_cast(double, v);
}

// ---- small math helpers --------------------------------------------------------
__device__ __forceinline__ float tmax(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ float tmin(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ double tmax(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ double tmin(double a, double b) { return __builtin_fmin(a, b); }
// Square roots enter only the wave-speed bounds.  FP32: hardware v_sqrt_f32
// (1 ulp).  FP64: v_rsq_f64 is only an approximation, so it seeds a Goldschmidt
// iteration on (g, h) -> (sqrt(x), 1/(2 sqrt(x))) plus one residual correction,
// accurate to about an ulp for x > 0; x == 0 (|v| of a resting state) gives 0.
__device__ __forceinline__ float tsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double tsqrt(double x) {
  double r = __builtin_amdgcn_rsq(x);
  double g = x * r, h = 0.5 * r;
  double e = __builtin_fma(-g, h, 0.5);
  g = __builtin_fma(g, e, g);
  h = __builtin_fma(h, e, h);
  e = __builtin_fma(-g, h, 0.5);
  g = __builtin_fma(g, e, g);
  h = __builtin_fma(h, e, h);
  double d = __builtin_fma(-g, g, x);
  return x > 0.0 ? __builtin_fma(d, h, g) : 0.0;
}
__device__ __forceinline__ float tabs(float x) { return __builtin_fabsf(x); }
__device__ __forceinline__ double tabs(double x) { return __builtin_fabs(x); }
// reciprocal: hardware v_rcp_f32 (1 ulp) in FP32; in FP64 v_rcp_f64 refined by
// two Newton steps r <- r + r(1 - x r) (about 1 ulp, half the cost of a division)
__device__ __forceinline__ double trcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = __builtin_fma(__builtin_fma(-x, r, 1.0), r, r);
  return __builtin_fma(__builtin_fma(-x, r, 1.0), r, r);
}
__device__ __forceinline__ float trcp(float x) { return __builtin_amdgcn_rcpf(x); }

// minmod of limiters.f90:1-21 with the third argument equal to the second:
// s*min(|a|,|b|) if a and b carry the same sign, else 0 -- which is the median
// of (a, b, 0): one v_med3_f32 in FP32.
__device__ __forceinline__ float minmod2(float a, float b) { return __builtin_amdgcn_fmed3f(a, b, 0.0f); }
__device__ __forceinline__ double minmod2(double a, double b) {
  return __builtin_fmax(__builtin_fmin(a, b), __builtin_fmin(__builtin_fmax(a, b), 0.0));
}

// a*b rounded to T before any later use: the product may not be fused into a
// following add (an FMA would cancel against an already-rounded value).
template <typename T>
__device__ __forceinline__ T mul_rounded(T a, T b) {
#pragma clang fp contract(off)
  return a * b;
}

template <typename T>
struct Vec4 { T r, mx, my, e; };

template <typename T>
__device__ __forceinline__ Vec4<T> operator+(Vec4<T> a, Vec4<T> b) { return {a.r + b.r, a.mx + b.mx, a.my + b.my, a.e + b.e}; }
template <typename T>
__device__ __forceinline__ Vec4<T> operator-(Vec4<T> a, Vec4<T> b) { return {a.r - b.r, a.mx - b.mx, a.my - b.my, a.e - b.e}; }
template <typename T>
__device__ __forceinline__ Vec4<T> scale(T s, Vec4<T> a) { return {s * a.r, s * a.mx, s * a.my, s * a.e}; }

template <typename T>
__device__ __forceinline__ Vec4<T> minmod2(Vec4<T> a, Vec4<T> b) {
  return {minmod2(a.r, b.r), minmod2(a.mx, b.mx), minmod2(a.my, b.my), minmod2(a.e, b.e)};
}
template <typename T>
__device__ __forceinline__ Vec4<T> from_left(Vec4<T> a) {
  return {lane_from_left(a.r), lane_from_left(a.mx), lane_from_left(a.my), lane_from_left(a.e)};
}
template <typename T>
__device__ __forceinline__ Vec4<T> from_right(Vec4<T> a) {
  return {lane_from_right(a.r), lane_from_right(a.mx), lane_from_right(a.my), lane_from_right(a.e)};
}

// ---- physical constants handed to the flux routines ------------------------------
template <typename T>
struct Phys {
  T gamma, gm1, half;   // gamma, gamma-1, 0.5
  T g;                  // |grad phi| per axis
  T ggm1, gfloor;       // gamma*(gamma-1) and gamma*1e-10 (sound speed of DEV sides)
};

// Face equilibrium at rest: density, pressure, energy p/(gamma-1) (the
// reference's compute_conservative of a resting state, benchmark_2d.f90:169),
// E+p, and gamma*p.
template <typename T>
struct FaceEq { T rho, p, e, ep, gp; };

enum { FLUX_RUSANOV = 0, FLUX_HLL = 1, FLUX_HLLC = 2 };
enum { FORM_PLAIN = 0, FORM_DEV = 1 };

// Per-side quantities.
template <typename T>
struct Side {
  Vec4<T> f;      // flux (PLAIN) or flux deviation (DEV) along the face normal
  T vn;           // signed normal velocity
  T c;            // sound speed with the reference's floors
  T vmag;         // |v| (Rusanov bound of the reference)
};

// PLAIN form: U = U_eq + dU in full, primitive and flux as compute_primitive /
// compute_flux (benchmark_2d.f90:153-156, 315-323).  One reciprocal per side
// (1/max(rho,1e-10)) serves both velocities and the sound speed; for physical
// states this differs from u(2)/w(1) by about one rounding.
template <typename T, int DIR>
__device__ __forceinline__ Side<T> side_plain(const Phys<T>& P, const FaceEq<T>& q, Vec4<T> d, Vec4<T>& U) {
  U = {q.rho + d.r, d.mx, d.my, q.e + d.e};
  T w0 = U.r;
  T inv = trcp(tmax(w0, T(1e-10)));
  T w1 = U.mx * inv;
  T w2 = U.my * inv;
  // the pressure is rounded once and shared by every use (see mul_rounded)
  T w3 = mul_rounded(P.gm1, U.e - P.half * w0 * (w1 * w1 + w2 * w2));
  Side<T> s;
  if (DIR == 0) {
    s.f = {w1 * U.r, w1 * U.mx + w3, w0 * w1 * w2, w1 * U.e + w1 * w3};
    s.vn = w1;
  } else {
    s.f = {U.r * w2, U.mx * w2, U.my * w2 + w3, w2 * U.e + w2 * w3};
    s.vn = w2;
  }
  s.c = tsqrt(P.gamma * tmax(w3, T(1e-10)) * inv);
  s.vmag = tsqrt(w1 * w1 + w2 * w2);
  return s;
}

// DEV form: with m_eq = 0, F(U_eq+dU) - F(U_eq) along x is
//   (dm_x, m_x u + dp, m_y u, (E+p) u),  dp = (gamma-1)(dE - |m|^2/(2 rho)),
// and symmetrically along y; no cancellation of O(1) terms occurs.  The sound
// speed takes gamma*p = gamma*p_eq + gamma(gamma-1)(dE - |m|^2/(2 rho)) in one
// fused step, floored at gamma*1e-10 (= gamma*max(p, 1e-10)).
template <typename T, int DIR>
__device__ __forceinline__ Side<T> side_dev(const Phys<T>& P, const FaceEq<T>& q, Vec4<T> d) {
  T rho = q.rho + d.r;
  T inv = trcp(tmax(rho, T(1e-10)));
  T u = d.mx * inv, v = d.my * inv;
  T de = d.e - P.half * (d.mx * u + d.my * v);
  T dp = P.gm1 * de;
  T ep = (q.ep + d.e) + dp;      // E + p
  Side<T> s;
  if (DIR == 0) {
    s.f = {d.mx, d.mx * u + dp, d.my * u, ep * u};
    s.vn = u;
  } else {
    s.f = {d.my, d.mx * v, d.my * v + dp, ep * v};
    s.vn = v;
  }
  s.c = tsqrt(tmax(P.ggm1 * de + q.gp, P.gfloor) * inv);
  s.vmag = tsqrt(u * u + v * v);
  return s;
}

// HLLC minus the equilibrium flux, in deviation form (DEV): the three-wave
// riemann_hllc of dg_with_source.f90:1318-1374 carried to 2D.  Wave speeds
// S_L = min(u_L,u_R) - max(c_L,c_R), S_R = max(u_L,u_R) + max(c_L,c_R) on the
// normal velocity; Lagrangian sound speeds D_L = rho_L (u_L - S_L),
// D_R = rho_R (S_R - u_R); the acoustic star state u* and p* - p_eq from the
// pressure deviations; the sampled state's flux with the tangential velocity of
// the contact's upwind side.  Every term of a star flux carries u*, and u* and
// p* - p_eq vanish exactly when both deviations do, so a resting atmosphere
// gives exactly zero.  q.e must hold the equilibrium energy.
template <typename T, int DIR>
__device__ __forceinline__ Vec4<T> hllc_dev(const Phys<T>& P, const FaceEq<T>& q, Vec4<T> dl, Vec4<T> dr) {
  struct HSide { T rho, un, ut, dp, c, e; Vec4<T> f; };
  auto side = [&](Vec4<T> d) {
    HSide s;
    s.rho = q.rho + d.r;
    const T inv = trcp(tmax(s.rho, T(1e-10)));
    const T u = d.mx * inv, v = d.my * inv;
    s.dp = P.gm1 * (d.e - P.half * (d.mx * u + d.my * v));
    const T ep = (q.ep + d.e) + s.dp;
    s.e = q.e + d.e;
    s.c = tsqrt(P.gamma * tmax(q.p + s.dp, T(1e-10)) * inv);
    if (DIR == 0) {
      s.un = u; s.ut = v;
      s.f = {d.mx, d.mx * u + s.dp, d.my * u, ep * u};
    } else {
      s.un = v; s.ut = u;
      s.f = {d.my, d.mx * v, d.my * v + s.dp, ep * v};
    }
    return s;
  };
  const HSide L = side(dl), R = side(dr);
  const T cm = tmax(L.c, R.c);
  const T sl = tmin(L.un, R.un) - cm, sr = tmax(L.un, R.un) + cm;
  const T DL = L.rho * (L.un - sl), DR = R.rho * (sr - R.un);
  const T inv = trcp(DL + DR);
  const T us = (DR * R.un + DL * L.un + (L.dp - R.dp)) * inv;
  const T dps = L.dp + (DL * (R.dp - L.dp) + DL * DR * (L.un - R.un)) * inv;   // p* - p_eq
  // star state of the contact's upwind side
  const bool lft = us > T(0);
  const T sk = lft ? sl : sr, rk = lft ? L.rho : R.rho, unk = lft ? L.un : R.un, utk = lft ? L.ut : R.ut;
  const T ek = lft ? L.e : R.e, pk = q.p + (lft ? L.dp : R.dp), ps = q.p + dps;
  const T den = trcp(sk - us);
  const T rs = rk * (sk - unk) * den;
  const T es = ((sk - unk) * ek - pk * unk + ps * us) * den;
  const T mf = rs * us;
  Vec4<T> fs;
  if (DIR == 0) fs = {mf, mf * us + dps, mf * utk, us * (es + ps)};
  else fs = {mf, mf * utk, mf * us + dps, us * (es + ps)};
  // sampled region: left state, a star state, or the right state
  const bool star = !(sl > T(0)) && (lft || sr > T(0));
  const Vec4<T>& side_f = sl > T(0) ? L.f : R.f;
  return star ? fs : side_f;
}

// Numerical flux minus the equilibrium flux at one face, given the deviations
// on the left (dl) and right (dr).  At dl == dr the Rusanov and HLL forms used
// here return the one-sided flux exactly (0.5*(f+f) = f; f + w*0 - q*0 = f), so
// the deviation is exactly zero at equilibrium.
template <typename T, int FLUX, int DIR, int FORM>
__device__ __forceinline__ Vec4<T> face_flux(const Phys<T>& P, const FaceEq<T>& q, Vec4<T> dl, Vec4<T> dr) {
  if constexpr (FLUX == FLUX_HLLC) {
    static_assert(FORM == FORM_DEV, "HLLC is implemented in deviation form");
    return hllc_dev<T, DIR>(P, q, dl, dr);
  }
  Side<T> L, R;
  Vec4<T> fe = {T(0), T(0), T(0), T(0)};
  Vec4<T> jl = dl, jr = dr;  // states entering the jump terms
  if (FORM == FORM_PLAIN) {
    L = side_plain<T, DIR>(P, q, dl, jl);
    R = side_plain<T, DIR>(P, q, dr, jr);
    // equilibrium flux through compute_flux of the resting face state: its
    // pressure is (gamma-1)*(E_eq - 0), equal to the one-sided flux at dU == 0
    // bit for bit.  It must be rounded before the final subtraction: fused
    // into fh - fe it would leave the product's rounding error behind.
    T pe = mul_rounded(P.gm1, q.e);
    if (DIR == 0) fe.mx = pe; else fe.my = pe;
  } else {
    L = side_dev<T, DIR>(P, q, dl);
    R = side_dev<T, DIR>(P, q, dr);
  }
  Vec4<T> fh;
  if (FLUX == FLUX_RUSANOV) {
    // compute_llflux: 0.5*(fR+fL) + 0.5*cmax*(uL-uR), cmax from |v|+c
    T cmax = tmax(L.vmag + L.c, R.vmag + R.c);
    T hc = P.half * cmax;
    fh = {P.half * (R.f.r + L.f.r) + hc * (jl.r - jr.r),
          P.half * (R.f.mx + L.f.mx) + hc * (jl.mx - jr.mx),
          P.half * (R.f.my + L.f.my) + hc * (jl.my - jr.my),
          P.half * (R.f.e + L.f.e) + hc * (jl.e - jr.e)};
  } else {
    // HLL, Davis bounds on the signed normal velocity, written as
    // F_L + w (F_R - F_L) - q (U_R - U_L), w = a-/(a+ + a-), q = a+ a-/(a+ + a-)
    T sl = tmin(L.vn - L.c, R.vn - R.c);
    T sr = tmax(L.vn + L.c, R.vn + R.c);
    T ap = tmax(T(0), sr), am = tmax(T(0), -sl);
    T inv = trcp(ap + am);
    T w = am * inv;
    if (FORM == FORM_DEV) {
      // the same flux with q = a+ w factored out: F_L + w ((F_R - F_L) - a+ (U_R - U_L))
      fh = {L.f.r + w * ((R.f.r - L.f.r) - ap * (jr.r - jl.r)),
            L.f.mx + w * ((R.f.mx - L.f.mx) - ap * (jr.mx - jl.mx)),
            L.f.my + w * ((R.f.my - L.f.my) - ap * (jr.my - jl.my)),
            L.f.e + w * ((R.f.e - L.f.e) - ap * (jr.e - jl.e))};
    } else {
      T qq = ap * am * inv;
      fh = {L.f.r + w * (R.f.r - L.f.r) - qq * (jr.r - jl.r),
            L.f.mx + w * (R.f.mx - L.f.mx) - qq * (jr.mx - jl.mx),
            L.f.my + w * (R.f.my - L.f.my) - qq * (jr.my - jl.my),
            L.f.e + w * (R.f.e - L.f.e) - qq * (jr.e - jl.e)};
    }
  }
  if (FORM == FORM_PLAIN) fh = fh - fe;
  return fh;
}

// Wave speed |v| + c of the cell state U_eq(centre) + dU (compute_speed,
// benchmark_2d.f90:283-295); ok is false for rho<=0, p<=0 or non-finite.
template <typename T>
__device__ __forceinline__ T cell_speed(const Phys<T>& P, T rho_eq, T e_eq, Vec4<T> d, bool& ok) {
  T rho = rho_eq + d.r;
  T inv = trcp(tmax(rho, T(1e-10)));
  T u = d.mx * inv, v = d.my * inv;
  T p = P.gm1 * ((e_eq + d.e) - P.half * rho * (u * u + v * v));
  T sp = tsqrt(u * u + v * v) + tsqrt(P.gamma * tmax(p, T(1e-10)) * inv);
  ok = (rho > T(0)) && (p > T(0)) && (sp < T(1e30));
  return sp;
}

}  // namespace fvwb
