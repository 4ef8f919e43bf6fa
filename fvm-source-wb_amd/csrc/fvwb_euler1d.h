// fvwb_euler1d.h — pointwise 1D Euler physics shared by the 1D finite-volume and
// DG kernels (host and device), in the reference's statement order.  It turns
// FMA contraction off for the rest of the including translation unit, so these
// functions, and the kernels that call them, round as written.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#pragma clang fp contract(off)

namespace fv1d {

// ---- pointwise physics (benchmark_1d.f90:115-198, 438-451; the same arithmetic in
// fvm.f90 and dg_with_source.f90:1167-1246, 1299-1316) --------------------------
__host__ __device__ inline void primitive(double gamma, const double u[3], double w[3]) {
  w[0] = u[0];
  w[1] = u[1] / w[0];
  w[2] = (gamma - 1.0) * (u[2] - 0.5 * w[0] * (w[1] * w[1]));
}

__host__ __device__ inline void conservative(double gamma, const double w[3], double u[3]) {
  u[0] = w[0];
  u[1] = w[0] * w[1];
  u[2] = w[2] / (gamma - 1.0) + 0.5 * w[0] * (w[1] * w[1]);
}

__host__ __device__ inline double speed(double gamma, const double u[3]) {
  double w[3];
  primitive(gamma, u, w);
  double cs = sqrt(gamma * fmax(w[2], 1e-10) / fmax(w[0], 1e-10));
  return fabs(w[1]) + cs;
}

__host__ __device__ inline void flux(double gamma, const double u[3], double f[3]) {
  double w[3];
  primitive(gamma, u, w);
  f[0] = w[1] * u[0];
  f[1] = w[1] * u[1] + w[2];
  f[2] = w[1] * u[2] + w[2] * w[1];
}

// compute_llflux(uleft, uright, f_left, f_right)
__host__ __device__ inline void llflux(double gamma, const double ul[3], const double ur[3], const double fl[3],
                                       const double fr[3], double out[3]) {
  double cmax = fmax(speed(gamma, ul), speed(gamma, ur));
  for (int k = 0; k < 3; ++k) out[k] = 0.5 * (fr[k] + fl[k]) - 0.5 * cmax * (ur[k] - ul[k]);
}

}  // namespace fv1d
