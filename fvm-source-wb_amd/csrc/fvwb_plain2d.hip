// fvwb_plain2d.hip — the plain, non-well-balanced 2D residual on the device.
//
// compute_update of benchmark_2d.f90:370-463: Rusanov (compute_llflux, :353-367)
// between the two cell states at every face, the Euler fluxes of compute_flux
// (:299-325), the pointwise source of get_source (:327-350, grad phi = (1,1)) and
// a frozen boundary ring.  FP64 on the host's AoS(4,nx,ny) layout, with the
// reference's statement order and no contraction, so the result is the
// reference's bit for bit.  The reference's last x face takes column ny (:418);
// that flux reaches only the ring, which is zeroed, so it is never formed here.
#include <hip/hip_runtime.h>

#pragma clang fp contract(off)

namespace fvwb {
namespace plain {

// compute_primitive (:145-157)
__device__ inline void prim(double gamma, const double* u, double* w) {
  w[0] = u[0];
  w[1] = u[1] / w[0];
  w[2] = u[2] / w[0];
  w[3] = (gamma - 1.0) * (u[3] - 0.5 * w[0] * (w[1] * w[1] + w[2] * w[2]));
}

// compute_speed (:283-295): |v| + c
__device__ inline double speed(double gamma, const double* u) {
  double w[4];
  prim(gamma, u, w);
  const double cs = sqrt(gamma * fmax(w[3], 1e-10) / fmax(w[0], 1e-10));
  return sqrt(w[1] * w[1] + w[2] * w[2]) + cs;
}

// compute_flux (:299-325) along x (dir 0) or y (dir 1)
__device__ inline void flux(double gamma, const double* u, int dir, double* f) {
  double w[4];
  prim(gamma, u, w);
  if (dir == 0) {
    f[0] = w[1] * u[0];
    f[1] = w[1] * u[1] + w[3];
    f[2] = w[0] * w[1] * w[2];
    f[3] = w[1] * u[3] + w[1] * w[3];
  } else {
    f[0] = u[0] * w[2];
    f[1] = u[1] * w[2];
    f[2] = u[2] * w[2] + w[3];
    f[3] = w[2] * u[3] + w[2] * w[3];
  }
}

// compute_llflux (:353-367) between cell states a (left/bottom) and b
__device__ inline void llflux(double gamma, const double* a, const double* b, int dir, double* out) {
  double fa[4], fb[4];
  flux(gamma, a, dir, fa);
  flux(gamma, b, dir, fb);
  const double cmax = fmax(speed(gamma, a), speed(gamma, b));
  for (int k = 0; k < 4; ++k) out[k] = 0.5 * (fb[k] + fa[k]) + 0.5 * cmax * (a[k] - b[k]);
}

__global__ void plain_update_kernel(const double* __restrict__ u, double* __restrict__ dudt, int nx, int ny,
                                    double gamma, double odx, double ody) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (long long)nx * ny) return;
  const int i = (int)(c % nx), j = (int)(c / nx);
  double* d = dudt + 4 * c;
  if (i == 0 || i == nx - 1 || j == 0 || j == ny - 1) {
    d[0] = d[1] = d[2] = d[3] = 0.0;
    return;
  }
  double uc[4], uw[4], ue[4], us[4], un[4];
  for (int k = 0; k < 4; ++k) {
    uc[k] = u[4 * c + k];
    uw[k] = u[4 * (c - 1) + k];
    ue[k] = u[4 * (c + 1) + k];
    us[k] = u[4 * (c - nx) + k];
    un[k] = u[4 * (c + nx) + k];
  }
  double Fl[4], Fr[4], Gb[4], Gt[4], w[4];
  llflux(gamma, uw, uc, 0, Fl);
  llflux(gamma, uc, ue, 0, Fr);
  llflux(gamma, us, uc, 1, Gb);
  llflux(gamma, uc, un, 1, Gt);
  prim(gamma, uc, w);
  const double s[4] = {0.0, -w[0] * 1.0, -w[0] * 1.0, -w[0] * (w[1] * 1.0 + w[2] * 1.0)};
  for (int k = 0; k < 4; ++k) d[k] = -(Fr[k] - Fl[k]) * odx - (Gt[k] - Gb[k]) * ody + s[k];
}

}  // namespace plain

hipError_t launch_plain_update(const double* u, double* dudt, int nx, int ny, double gamma, double boxlen_x,
                               double boxlen_y, hipStream_t s) {
  const double dx = boxlen_x / (double)nx, dy = boxlen_y / (double)ny;
  const long long n = (long long)nx * ny;
  hipLaunchKernelGGL(plain::plain_update_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, u, dudt, nx, ny,
                     gamma, 1.0 / dx, 1.0 / dy);
  return hipGetLastError();
}

}  // namespace fvwb
