// fvwb_internal.h — structures shared by the kernel and ABI translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fvwb {

// Device-side step control of one instance.  Each step n uses slot p = n & 1;
// its last stage accumulates max(|v|+c) of the new state into cmax[1-p] and
// writes t[1-p], so no host synchronisation is needed between steps.
// Atomics on one cache line serialise, so every slot has a 128-byte line of its own.
struct alignas(128) SpeedSlot {
  unsigned long long v;           // bit pattern of a non-negative double
  unsigned long long pad[15];
};
struct Ctrl {
  SpeedSlot cmax[2][8];           // one slot per XCD
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
  long long plane;    // element stride between the variables (pitch * ny for separate planes)
  int pitch, nx, ny;  // pitch: element stride between grid rows
  int rows;           // rows per wave strip
  int ntx, ntiles;    // strips along x, total strips
  int parity, first, last;
  double t_end, h, cfl;
  double gamma, g, p_over_rho, inv_gm1, odx, ody;
  double alpha, beta;
};

// Per-instance description of a batch member (device array).
struct InstDesc {
  long long off;      // element offset of the instance's 4-plane block in a pool buffer
  long long plane;    // element stride between the variables (P: rows side by side)
  int pitch, nx, ny;  // pitch: element stride between grid rows (4 P)
  int rows, ntx, ntiles;
  int tx_off, ty_off; // offsets of the x tables (rxf[nx+1], rxc[nx]) and y tables (eyc[ny], eyf[ny+1])
  double h, cfl, gamma, g, p_over_rho, inv_gm1, odx, ody;
};

// Kernel arguments of one batched stage.
template <typename T>
struct BatchArgs {
  const InstDesc* inst;
  const int* tile_start;  // prefix sums of tiles per instance, [ninst]
  int ninst, ntiles;
  const T* in;
  const T* un;
  T* out;
  const double* tabx;
  const double* taby;
  Ctrl* ctrl;              // one per instance
  int parity, first, last;
  double t_end, alpha, beta;
};

// Launch configuration chosen by the tuner.
struct Tune {
  int rows;  // rows per wave strip
  int cpl;   // FP32 columns per lane (2 or 4)
};

}  // namespace fvwb
