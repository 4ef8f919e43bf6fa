// fvwb_internal.h — structures shared by the kernel and ABI translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fvwb {

// Device-side step control of one instance.  Each step n uses slot p = n & 1;
// its last stage accumulates max(|v|+c) of the new state into cmax[1-p] and
// writes t[1-p], so no host synchronisation is needed between steps.
struct Ctrl {
  unsigned long long cmax[2][8];  // bit patterns of non-negative doubles, one per XCD
  double t[2];
  double dt_last;
  int32_t steps_done;
  uint32_t err;                   // bit 0: rho<=0, p<=0 or non-finite seen
};

// Kernel arguments of one 2D FV stage (one instance).
template <typename T>
struct StageArgs {
  const T* in;        // deviation planes read with halo
  const T* un;        // deviation planes at t^n (stage modes with alpha != 0)
  T* out;             // output planes (may alias un, never in)
  const double* rxf;  // rho0*exp(-k x) at x faces   [nx+1]
  const double* rxc;  // rho0*exp(-k x) at centres   [nx]
  const double* eyc;  // exp(-k y) at centres        [ny]
  const double* eyf;  // exp(-k y) at y faces        [ny+1]
  Ctrl* ctrl;
  long long plane;    // elements per variable plane
  int pitch, nx, ny;
  int rows;           // rows per wave strip
  int ntx, ntiles;    // strips along x, total strips
  int parity, first, last;
  double t_end, h, cfl;
  double gamma, g, p_over_rho, inv_gm1, odx, ody;
  double alpha, beta;
};

// Launch configuration chosen by the tuner.
struct Tune {
  int rows;  // rows per wave strip
};

}  // namespace fvwb
