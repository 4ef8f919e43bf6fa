// fvwb_fv1d.hip — 1D well-balanced Euler with gravity (benchmark_1d.f90) on MI355X.
//
// Device paths:
//  * fv1d_ref_update_kernel: the drop-in for compute_update (benchmark_1d.f90:
//    263-377), one thread per cell, global memory, FP64 rounding statement by
//    statement (no contraction, IEEE division and square root), so it returns
//    the reference's dudt bit for bit given the same face equilibrium table.
//  * fv1d_ref_update_sr_kernel: the drop-in for compute_update_sr (:553-747),
//    the hydrostatic-reconstruction solver 'WB1', in the same style;
//  * fv1d_plain_update_kernel: the plain residuals, compute_update_fvm
//    (:454-549, solver 'FVM') and compute_update of the stand-alone fvm.f90.
//  * fv1d_run_kernel: the configured scheme for a batch of independent 1D
//    problems.  A 1D problem of up to 2048 cells fits in one CU's LDS, so one
//    workgroup owns one instance for the whole run: the state lives in
//    registers (each lane owns cells i = lane + k*256), the stage input and the
//    face fluxes are staged in LDS, the CFL maximum is a wave reduction plus a
//    4-entry LDS combine, and every time step of the run happens inside one
//    launch with no host round trip.  HBM is touched only at the start and end
//    of the run.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/fvwb.h"
#include "fvwb_checkpoint.h"
#include "fvwb_euler1d.h"

// Every expression in this file rounds as written (the oracle and the
// reference are built without FMA contraction).
#pragma clang fp contract(off)

namespace fv1d {

constexpr int BS = 256;        // lanes per workgroup (4 waves)
constexpr int MAX_NX = 2048;   // cells per instance in the LDS-resident kernel

struct Inst {
  long long off;     // first cell of this instance in the state pools (planes of 3*nx)
  long long toff;    // first face of this instance in the face-table pool (3*(nx+1))
  int nx;
  int pad;
  double dx, odx, cfl, g, gamma;
};

struct Clock {
  double t;
  double dt_last;
  int32_t steps;
  uint32_t err;     // bit 0: rho<=0, p<=0 or non-finite seen; bit 1: dt not positive
};

// limiters.f90:1-21 minmod(x, y, y): the common-sign minimum magnitude
__device__ inline double minmod2(double x, double y) {
  const bool sx = signbit(x), sy = signbit(y);
  if (sx != sy) return 0.0;
  const double m = fmin(fabs(x), fabs(y));
  return sx ? -m : m;
}

// Numerical flux minus equilibrium flux at one face of the configured scheme:
// face equilibrium qe (conservative) with flux fe, cell deviations am, a | b, bp
// around the face (am, bp used by MUSCL only).
template <int RECON>
__device__ inline void face_G(double gamma, const double qe[3], const double fe[3], const double am[3],
                              const double a[3], const double b[3], const double bp[3], double out[3]) {
  double dl[3], dr[3];
  for (int v = 0; v < 3; ++v) {
    if (RECON == FVWB_RECON_MUSCL_MINMOD) {
      const double sa = minmod2(a[v] - am[v], b[v] - a[v]);
      const double sb = minmod2(b[v] - a[v], bp[v] - b[v]);
      dl[v] = a[v] + 0.5 * sa;
      dr[v] = b[v] - 0.5 * sb;
    } else {
      dl[v] = a[v];
      dr[v] = b[v];
    }
  }
  double ul[3], ur[3], fl[3], fr[3], fh[3];
  for (int v = 0; v < 3; ++v) { ul[v] = qe[v] + dl[v]; ur[v] = qe[v] + dr[v]; }
  flux(gamma, ul, fl);
  flux(gamma, ur, fr);
  llflux(gamma, ul, ur, fl, fr, fh);
  for (int v = 0; v < 3; ++v) out[v] = fh[v] - fe[v];
}

// ---- drop-in: benchmark_1d.f90:263-377 ------------------------------------------
// u, w_eq, dudt: AoS(3,nx); ueqf: conservative face equilibrium AoS(3,nx+1).
__global__ __launch_bounds__(256) void fv1d_ref_update_kernel(const double* __restrict__ u,
                                                              const double* __restrict__ w_eq,
                                                              const double* __restrict__ ueqf,
                                                              double* __restrict__ dudt, int nx, double gamma,
                                                              double dx, double oneoverdx, double delta_s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx) return;
  // boundary cells take their neighbour's value (:374-375)
  const int c = i == 0 ? 1 : (i == nx - 1 ? nx - 2 : i);
  double dl[3], dc[3], dr[3], uec[3];
  {
    double ue[3];
    conservative(gamma, &w_eq[3 * (c - 1)], ue);
    for (int k = 0; k < 3; ++k) dl[k] = u[3 * (c - 1) + k] - ue[k];
    conservative(gamma, &w_eq[3 * c], uec);
    for (int k = 0; k < 3; ++k) dc[k] = u[3 * c + k] - uec[k];
    conservative(gamma, &w_eq[3 * (c + 1)], ue);
    for (int k = 0; k < 3; ++k) dr[k] = u[3 * (c + 1) + k] - ue[k];
  }
  // faces c (between c-1 and c) and c+1 (between c and c+1)
  double FR[2][3];
  for (int s = 0; s < 2; ++s) {
    const double* da = s == 0 ? dl : dc;   // cell on the left of the face
    const double* db = s == 0 ? dc : dr;   // cell on the right of the face
    const int f = c + s;
    double a[3], b[3], fa[3], fb[3];
    for (int k = 0; k < 3; ++k) {
      a[k] = da[k] + ueqf[3 * f + k];   // u_right of the left cell
      b[k] = db[k] + ueqf[3 * f + k];   // u_left of the right cell
    }
    flux(gamma, a, fa);
    flux(gamma, b, fb);
    llflux(gamma, a, b, fa, fb, FR[s]);
  }
  double FE[2][3];
  flux(gamma, &ueqf[3 * c], FE[0]);
  flux(gamma, &ueqf[3 * (c + 1)], FE[1]);
  // get_source (:379-405) evaluated on the conservative states
  const double xm = ((double)(c - 1 + 1) - 0.5) * dx;
  const double xp = ((double)(c + 1 + 1) - 0.5) * dx;
  const double* uc = &u[3 * c];
  double s[3], se[3];
  s[0] = 0;
  s[1] = -(uc[0] * 1 * (xp - xm) / (2 * delta_s));
  s[2] = -(uc[0] * uc[1] * 1 * (xp - xm) / (2 * delta_s));
  se[0] = 0;
  se[1] = -(uec[0] * 1 * (xp - xm) / (2 * delta_s));
  se[2] = -(uec[0] * uec[1] * 1 * (xp - xm) / (2 * delta_s));
  for (int k = 0; k < 3; ++k)
    dudt[3 * i + k] = -((FR[1][k] - FR[0][k]) * oneoverdx) + s[k] + (FE[1][k] - FE[0][k]) * oneoverdx - se[k];
}

// ---- drop-in: compute_update_sr, benchmark_1d.f90:553-747 ('WB1') ---------------
// Hydrostatic reconstruction of one cell's face states (:596-629): entropy
// K = max(p,1e-5)/max(rho,1e-5)^gamma and enthalpy h = gamma/(gamma-1) p/rho are
// carried from the centre to each face along phi = x, the velocity is the
// cell's.  Cell i (zero based) has centre (i+0.5)*dx and faces i*dx, (i+1)*dx.
// The 1e-5 floors are FP32 literals in the reference.  pow is the device
// library's, so this path agrees with the reference to a few ulps, not bit for
// bit.
__device__ inline void sr_states(double gamma, double dx, int i, const double* __restrict__ u, double w[3],
                                 double wl[3], double wr[3]) {
  primitive(gamma, &u[3 * i], w);
  const double e5 = (double)1e-5f;
  const double h = fmax(w[2], e5) / fmax(e5, w[0]) * (1 + 1.0 / (gamma - 1));
  const double xc = ((double)(i + 1) - 0.5) * dx;
  const double h0[2] = {h + xc - (double)i * dx, h + xc - (double)(i + 1) * dx};
  const double kap = fmax(e5, w[2]) / pow(fmax(e5, w[0]), gamma);
  const double ik = pow(1.0 / kap, 1 / (gamma - 1));
  double* o[2] = {wl, wr};
  for (int s = 0; s < 2; ++s) {
    o[s][0] = pow((1.0 / kap) * (gamma - 1) / gamma * h0[s], 1 / (gamma - 1));
    o[s][1] = w[1];
    o[s][2] = ik * pow((gamma - 1) / gamma * h0[s], gamma / (gamma - 1));
  }
}

// One thread per cell.  Interior faces only carry Rusanov fluxes (:637-665: the
// face loop runs over 2..nx, so its boundary branches never execute), the
// source is get_source_rg (:409-434) and the boundary cells copy their
// neighbours (:741-742).
__global__ __launch_bounds__(256) void fv1d_ref_update_sr_kernel(const double* __restrict__ u,
                                                                 double* __restrict__ dudt, int nx, double gamma,
                                                                 double dx, double oneoverdx, double delta) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx) return;
  const int c = i == 0 ? 1 : (i == nx - 1 ? nx - 2 : i);
  double w[3], wl[3], wr[3], wa[3], wb[3], t[3];
  sr_states(gamma, dx, c, u, w, wl, wr);
  double ul[3], ur[3], fl[3], fr[3];
  conservative(gamma, wl, ul);
  conservative(gamma, wr, ur);
  flux(gamma, ul, fl);
  flux(gamma, ur, fr);
  double F0[3], F1[3];
  {
    double ua[3], fa[3];
    sr_states(gamma, dx, c - 1, u, t, wa, wb);   // right face state of cell c-1
    conservative(gamma, wb, ua);
    flux(gamma, ua, fa);
    llflux(gamma, ua, ul, fa, fl, F0);
    sr_states(gamma, dx, c + 1, u, t, wa, wb);   // left face state of cell c+1
    conservative(gamma, wa, ua);
    flux(gamma, ua, fa);
    llflux(gamma, ur, ua, fr, fa, F1);
  }
  const double xm = ((double)c - 0.5) * dx;
  const double xp = ((double)(c + 2) - 0.5) * dx;
  const double s[3] = {0.0, (wr[2] - wl[2]) / delta, -(w[0] * w[1] * 1 * (xp - xm) / (2 * delta))};
  for (int k = 0; k < 3; ++k) dudt[3 * i + k] = -((F1[k] - F0[k]) * oneoverdx) + s[k];
}

// ---- drop-ins: the plain (not well-balanced) 1D residuals -----------------------
// PROGRAM == false: compute_update_fvm (benchmark_1d.f90:454-549, solver 'FVM'):
//   Rusanov between the cell states, get_source (:379-405) on the conservative
//   state with delta = 1/nx, boundary cells copy their neighbours (:546-547).
// PROGRAM == true: compute_update of the stand-alone fvm.f90 (:188-251): the
//   end cells' outer neighbours wrap (periodic) or repeat the cell (zero
//   gradient), the optional source uses the primitive state with x- and x+
//   clamped to the cell's own centre at the two ends (:232-243) and delta = dx;
//   no boundary copy.
// One thread per cell, both faces of the cell computed by that thread.
template <bool PROGRAM>
__global__ __launch_bounds__(256) void fv1d_plain_update_kernel(const double* __restrict__ u,
                                                                double* __restrict__ dudt, int nx, double gamma,
                                                                double dx, double oneoverdx, double delta,
                                                                int periodic, int source) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx) return;
  int c = i, im = i - 1, ip = i + 1;
  if (!PROGRAM) {
    c = i == 0 ? 1 : (i == nx - 1 ? nx - 2 : i);
    im = c - 1;
    ip = c + 1;
  } else {
    if (i == 0) im = periodic ? nx - 1 : 0;
    if (i == nx - 1) ip = periodic ? 0 : nx - 1;
  }
  const double* um = &u[3 * im];
  const double* uc = &u[3 * c];
  const double* up = &u[3 * ip];
  double fm[3], fc[3], fp[3], FL[3], FR[3];
  flux(gamma, um, fm);
  flux(gamma, uc, fc);
  flux(gamma, up, fp);
  llflux(gamma, um, uc, fm, fc, FL);
  llflux(gamma, uc, up, fc, fp, FR);
  double s[3] = {0.0, 0.0, 0.0};
  if (!PROGRAM) {
    const double xm = ((double)c - 0.5) * dx;
    const double xp = ((double)(c + 2) - 0.5) * dx;
    s[1] = -(uc[0] * 1 * (xp - xm) / (2 * delta));
    s[2] = -(uc[0] * uc[1] * 1 * (xp - xm) / (2 * delta));
    for (int k = 0; k < 3; ++k) dudt[3 * i + k] = -((FR[k] - FL[k]) * oneoverdx) + s[k];
  } else {
    if (source == 2) {
      const double xm = (i == 0 ? 0.5 : (double)i - 0.5) * dx;
      const double xp = (i == nx - 1 ? (double)nx - 0.5 : (double)(i + 2) - 0.5) * dx;
      double w[3];
      primitive(gamma, uc, w);
      s[1] = -(w[0] * 1 * (xp - xm) / (2 * delta));
      s[2] = -(w[0] * w[1] * 1 * (xp - xm) / (2 * delta));
    }
    for (int k = 0; k < 3; ++k) dudt[3 * i + k] = -(oneoverdx * (FR[k] - FL[k])) + s[k];
  }
}

// compute_max_speed (:157-169) over AoS(3,n)
__global__ __launch_bounds__(256) void fv1d_max_speed_kernel(const double* __restrict__ u, int n, double gamma,
                                                             unsigned long long* __restrict__ out) {
  double m = 0.0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
    m = fmax(m, speed(gamma, &u[3 * c]));
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

// ---- configured scheme: one workgroup runs one instance ------------------------------
// LDS: D[3][nx+4] (stage deviation with two ghost cells each side), G[3][nx+1]
// (face flux minus equilibrium flux), wmax[4].
template <int RECON, int BC, int INTEG, int C>
__global__ __launch_bounds__(BS) void fv1d_run_kernel(const Inst* __restrict__ insts, double* __restrict__ dev,
                                                      const double* __restrict__ ueqc_pool,
                                                      const double* __restrict__ ueqf_pool, Clock* __restrict__ clk,
                                                      int nsteps, double t_end) {
  extern __shared__ double lds[];
  const Inst I = insts[blockIdx.x];
  const int nx = I.nx, ld = nx + 4;
  double* D = lds;                 // D[v*ld + i + 2]
  double* G = lds + 3 * ld;        // G[v*(nx+1) + f]
  double* wmax = G + 3 * (nx + 1);
  const int lane = threadIdx.x;
  double* st = dev + 3 * I.off;                // planes: st[v*nx + i]
  const double* ueqc = ueqc_pool + 3 * I.off;  // planes
  const double* ueqf = ueqf_pool + 3 * I.toff; // planes: ueqf[v*(nx+1) + f]
  const double gamma = I.gamma;

  // own cells and faces
  double dn[C][3], ds[C][3], ue[C][3];
  double fe[C + 1][3], qe[C + 1][3];   // F(U_eq) and U_eq at own faces
#pragma unroll
  for (int k = 0; k < C; ++k) {
    const int i = lane + k * BS;
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      dn[k][v] = i < nx ? st[v * nx + i] : 0.0;
      ue[k][v] = i < nx ? ueqc[v * nx + i] : 1.0;
      ds[k][v] = dn[k][v];
    }
  }
#pragma unroll
  for (int k = 0; k <= C; ++k) {
    const int f = lane + k * BS;
    for (int v = 0; v < 3; ++v) qe[k][v] = f <= nx ? ueqf[v * (nx + 1) + f] : (v == 0 ? 1.0 : (v == 1 ? 0.0 : 2.5));
    flux(gamma, qe[k], fe[k]);
  }
  Clock ck = clk[blockIdx.x];
  const int wave = lane >> 6;

  for (int n = 0; n < nsteps && ck.t < t_end && ck.err == 0; ++n) {
    // ---- dt: max |v| + c of U_eq + dU over the instance
    double m = 0.0;
    bool bad = false;
#pragma unroll
    for (int k = 0; k < C; ++k) {
      if (lane + k * BS < nx) {
        double s[3] = {ue[k][0] + dn[k][0], ue[k][1] + dn[k][1], ue[k][2] + dn[k][2]};
        double w[3];
        primitive(gamma, s, w);
        bad |= !(w[0] > 0.0) || !(w[2] > 0.0) || !isfinite(s[0] + s[1] + s[2]);
        m = fmax(m, speed(gamma, s));
      }
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    const int anybad = __syncthreads_or(bad ? 1 : 0);
    if (lane == 0 || lane == 64 || lane == 128 || lane == 192) wmax[wave] = m;
    __syncthreads();
    if (anybad) {
      ck.err |= 1u;
      break;
    }
    const double cmax = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
    double dt = I.cfl * I.dx / cmax;
    if (ck.t + dt > t_end) dt = t_end - ck.t;
    if (!(dt > 0.0)) {      // uniform across the workgroup: every lane computed the same dt
      ck.err |= 2u;
      break;
    }

    const int nstage = INTEG == FVWB_SSPRK3 ? 3 : 2;
    for (int stage = 0; stage < nstage; ++stage) {
      // stage input -> LDS; the ghost cells are written by the owners of the
      // cells they copy (walls: mirrored, momentum negated; zero gradient:
      // copies of the boundary cell)
#pragma unroll
      for (int k = 0; k < C; ++k) {
        const int i = lane + k * BS;
        if (i < nx) {
          for (int v = 0; v < 3; ++v) D[v * ld + i + 2] = ds[k][v];
          if (BC == FVWB1D_BC_WALL) {
            int gi = -100;
            if (i <= 1) gi = -1 - i;                       // ghosts -1, -2 from cells 0, 1
            else if (i >= nx - 2) gi = 2 * nx - 1 - i;     // ghosts nx, nx+1 from cells nx-1, nx-2
            if (gi != -100) {
              D[0 * ld + gi + 2] = ds[k][0];
              D[1 * ld + gi + 2] = -ds[k][1];
              D[2 * ld + gi + 2] = ds[k][2];
            }
          } else if (i == 0 || i == nx - 1) {
            const int g0 = i == 0 ? 0 : nx + 2, g1 = i == 0 ? 1 : nx + 3;
            for (int v = 0; v < 3; ++v) {
              D[v * ld + g0] = ds[k][v];
              D[v * ld + g1] = ds[k][v];
            }
          }
        }
      }
      __syncthreads();
      // face fluxes: face f lies between cells f-1 and f
#pragma unroll
      for (int k = 0; k <= C; ++k) {
        const int f = lane + k * BS;
        if (f <= nx) {
          double am[3], a[3], b[3], bp[3], gf[3];
          for (int v = 0; v < 3; ++v) {
            am[v] = D[v * ld + f];
            a[v] = D[v * ld + f + 1];
            b[v] = D[v * ld + f + 2];
            bp[v] = D[v * ld + f + 3];
          }
          face_G<RECON>(gamma, qe[k], fe[k], am, a, b, bp, gf);
          for (int v = 0; v < 3; ++v) G[v * (nx + 1) + f] = gf[v];
        }
      }
      __syncthreads();
      // residual and stage combination
#pragma unroll
      for (int k = 0; k < C; ++k) {
        const int i = lane + k * BS;
        if (i < nx) {
          double L[3];
          const double s1 = -I.g * ds[k][0], s2 = -I.g * ds[k][1];
          L[0] = -(G[0 * (nx + 1) + i + 1] - G[0 * (nx + 1) + i]) * I.odx + 0.0;
          L[1] = -(G[1 * (nx + 1) + i + 1] - G[1 * (nx + 1) + i]) * I.odx + s1;
          L[2] = -(G[2 * (nx + 1) + i + 1] - G[2 * (nx + 1) + i]) * I.odx + s2;
          for (int v = 0; v < 3; ++v) {
            if (stage == 0) {
              ds[k][v] = dn[k][v] + dt * L[v];
            } else if (INTEG == FVWB_SSPRK3) {
              if (stage == 1) ds[k][v] = 0.75 * dn[k][v] + 0.25 * ds[k][v] + 0.25 * dt * L[v];
              else ds[k][v] = (1.0 / 3.0) * dn[k][v] + (2.0 / 3.0) * ds[k][v] + (2.0 / 3.0) * dt * L[v];
            } else {
              ds[k][v] = 0.5 * dn[k][v] + 0.5 * ds[k][v] + 0.5 * dt * L[v];
            }
          }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < C; ++k)
      for (int v = 0; v < 3; ++v) dn[k][v] = ds[k][v];
    ck.t = ck.t + dt;
    ck.dt_last = dt;
    ck.steps += 1;
  }
#pragma unroll
  for (int k = 0; k < C; ++k) {
    const int i = lane + k * BS;
    if (i < nx)
      for (int v = 0; v < 3; ++v) st[v * nx + i] = dn[k][v];
  }
  if (lane == 0) clk[blockIdx.x] = ck;
}

// Residual of the configured scheme on deviation planes d (one instance), one
// lane per cell; ghost cells by index (walls mirror, zero gradient clamps).
template <int BC>
__device__ inline void dev_at(const double* d, int nx, int i, double out[3]) {
  int src = i;
  bool mirror = false;
  if (i < 0) { if (BC == FVWB1D_BC_WALL) { src = -1 - i; mirror = true; } else src = 0; }
  else if (i >= nx) { if (BC == FVWB1D_BC_WALL) { src = 2 * nx - 1 - i; mirror = true; } else src = nx - 1; }
  out[0] = d[src];
  out[1] = mirror ? -d[nx + src] : d[nx + src];
  out[2] = d[2 * nx + src];
}

template <int RECON, int BC>
__global__ __launch_bounds__(256) void fv1d_residual_kernel(const double* __restrict__ d,
                                                            const double* __restrict__ ueqf, double* __restrict__ L,
                                                            int nx, double gamma, double odx, double g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx) return;
  double cells[5][3];
  for (int q = 0; q < 5; ++q) dev_at<BC>(d, nx, i - 2 + q, cells[q]);
  double G[2][3];
  for (int s = 0; s < 2; ++s) {
    const int f = i + s;
    double qe[3], fe[3];
    for (int v = 0; v < 3; ++v) qe[v] = ueqf[v * (nx + 1) + f];
    flux(gamma, qe, fe);
    face_G<RECON>(gamma, qe, fe, cells[s], cells[s + 1], cells[s + 2], cells[s + 3], G[s]);
  }
  const double s1 = -g * cells[2][0], s2 = -g * cells[2][1];
  L[3 * i + 0] = -(G[1][0] - G[0][0]) * odx + 0.0;
  L[3 * i + 1] = -(G[1][1] - G[0][1]) * odx + s1;
  L[3 * i + 2] = -(G[1][2] - G[0][2]) * odx + s2;
}

// u (AoS(3,nx)) = U_eq(centre) + dU (planes)
__global__ void fv1d_to_aos_kernel(const double* __restrict__ d, const double* __restrict__ ueqc,
                                   double* __restrict__ u, int nx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx) return;
  for (int v = 0; v < 3; ++v) u[3 * i + v] = ueqc[v * nx + i] + d[v * nx + i];
}

}  // namespace fv1d

using namespace fv1d;

struct fvwb1d {
  std::vector<fvwb1d_params> p;
  std::vector<Inst> inst;
  int ninst = 0, cmax_per_lane = 1;
  long long cells = 0, faces = 0;
  hipStream_t stream = nullptr;
  Inst* d_inst = nullptr;
  double* d_dev = nullptr;     // dU planes, all instances
  double* d_ueqc = nullptr;    // conservative centre equilibrium planes (from w_eq)
  double* d_ueqf = nullptr;    // conservative face equilibrium planes (closed form)
  double* d_reff = nullptr;    // AoS(3,nx+1) face equilibrium of instance 0 as the reference computes it
  double* d_stage = nullptr;   // staging for host arrays: 3 arrays of 3*max(nx) doubles
  Clock* d_clk = nullptr;
  unsigned long long* d_red = nullptr;
  std::vector<char> ready;
  size_t lds_bytes = 0;
  std::string err;
};

static thread_local std::string g_err1d;

static int fail1d(fvwb1d* h, int code, const std::string& msg) {
  if (h) h->err = msg; else g_err1d = msg;
  return code;
}

#define HIPC(h, expr)                                                           \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess) return fail1d((h), FVWB_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// closed-form equilibrium, primitive, of p at x (general g; the reference's
// get_equilibrium_solution has g = 1)
static void equilibrium_host(const fvwb1d_params& p, double x, double w[3]) {
  const double gamma = p.gamma;
  if (p.nequilibrium == 3) {
    double base = 1 - ((gamma - 1) / gamma) * p.g * x;
    w[0] = std::pow(base, 1 / (gamma - 1));
    w[1] = 0;
    w[2] = std::pow(base, gamma / (gamma - 1));
  } else {
    w[0] = std::exp(-(p.g * x));
    w[1] = 0;
    w[2] = std::exp(-(p.g * x));
  }
}

// the reference's own face equilibrium (benchmark_1d.f90:134-155, 297-300)
static void equilibrium_ref(const fvwb1d_params& p, double x, double w[3]) {
  const double gamma = p.gamma;
  if (p.nequilibrium == 3) {
    double base = 1 - ((gamma - 1) / gamma) * 1 * x;
    w[0] = std::pow(base, 1 / (gamma - 1));
    w[1] = 0;
    w[2] = std::pow(base, gamma / (gamma - 1));
  } else {
    w[0] = std::exp(-x);
    w[1] = 0;
    w[2] = std::exp(-x);
  }
}

extern "C" int fvwb1d_params_reference(fvwb1d_params* p, int32_t nx, int32_t nequilibrium) {
  if (!p || nequilibrium < 1 || nequilibrium > 3) return FVWB_ERR_ARG;
  std::memset(p, 0, sizeof *p);
  p->nvar = 3; p->nx = nx;
  p->bc = FVWB1D_BC_ZERO_GRADIENT;                 // parameters.f90:12
  p->recon = FVWB_RECON_FIRST_ORDER; p->integrator = FVWB_RK2_HEUN;
  p->nequilibrium = nequilibrium;
  p->gamma = (double)1.4f;                         // parameters.f90:17
  p->boxlen = 1.0; p->cfl = 0.5; p->g = 1.0;
  return FVWB_OK;
}

static bool validate1d(const fvwb1d_params& p, std::string& why) {
  if (p.nvar != 3) { why = "nvar must be 3"; return false; }
  if (p.nx < 4) { why = "nx must be >= 4"; return false; }
  if (p.bc < 1 || p.bc > 4) { why = "bc must be a FVWB1D_BC_* code"; return false; }
  if (p.recon != FVWB_RECON_FIRST_ORDER && p.recon != FVWB_RECON_MUSCL_MINMOD) { why = "bad recon"; return false; }
  if (p.integrator != FVWB_RK2_HEUN && p.integrator != FVWB_SSPRK3) { why = "bad integrator"; return false; }
  if (p.nequilibrium < 1 || p.nequilibrium > 3) { why = "nequilibrium must be 1, 2 or 3"; return false; }
  if (!(p.gamma > 1.0) || !(p.boxlen > 0.0) || !(p.cfl > 0.0) || !(p.g >= 0.0) || !std::isfinite(p.g)) {
    why = "gamma > 1, boxlen > 0, cfl > 0 and g >= 0 required";
    return false;
  }
  if (p.nequilibrium == 3 && !(1 - ((p.gamma - 1) / p.gamma) * p.g * p.boxlen > 0)) {
    why = "polytropic equilibrium reaches zero density inside the box";
    return false;
  }
  return true;
}

extern "C" int fvwb1d_destroy(fvwb1d_handle h) {
  if (!h) return FVWB_OK;
  void* ps[] = {h->d_inst, h->d_dev, h->d_ueqc, h->d_ueqf, h->d_reff, h->d_stage, h->d_clk, h->d_red};
  for (void* q : ps)
    if (q) (void)hipFree(q);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return FVWB_OK;
}

extern "C" int fvwb1d_create(fvwb1d_handle* out, int32_t ninst, const fvwb1d_params* per_inst) {
  if (!out) return fail1d(nullptr, FVWB_ERR_ARG, "handle pointer is NULL");
  *out = nullptr;
  if (ninst < 1 || !per_inst) return fail1d(nullptr, FVWB_ERR_ARG, "ninst >= 1 and params required");
  std::string why;
  for (int b = 0; b < ninst; ++b) {
    if (!validate1d(per_inst[b], why)) return fail1d(nullptr, FVWB_ERR_ARG, "instance " + std::to_string(b) + ": " + why);
    const fvwb1d_params &a = per_inst[b], &z = per_inst[0];
    if (a.recon != z.recon || a.integrator != z.integrator || a.bc != z.bc)
      return fail1d(nullptr, FVWB_ERR_ARG, "instances must share recon, integrator and bc");
  }
  fvwb1d* h = new (std::nothrow) fvwb1d();
  if (!h) return fail1d(nullptr, FVWB_ERR_OOM, "host allocation failed");
  h->ninst = ninst;
  h->p.assign(per_inst, per_inst + ninst);
  int maxnx = 0;
  for (int b = 0; b < ninst; ++b) {
    Inst I{};
    I.off = h->cells; I.toff = h->faces; I.nx = per_inst[b].nx;
    I.dx = per_inst[b].boxlen / (double)I.nx; I.odx = 1.0 / I.dx;
    I.cfl = per_inst[b].cfl; I.g = per_inst[b].g; I.gamma = per_inst[b].gamma;
    h->inst.push_back(I);
    h->cells += I.nx; h->faces += I.nx + 1;
    if (I.nx > maxnx) maxnx = I.nx;
  }
  h->cmax_per_lane = 1;
  while (h->cmax_per_lane * BS < maxnx) h->cmax_per_lane *= 2;
  h->lds_bytes = sizeof(double) * (3 * (size_t)(maxnx + 4) + 3 * (size_t)(maxnx + 1) + 4);
  h->ready.assign(ninst, 0);
  auto bail = [&](int rc) { g_err1d = h->err; fvwb1d_destroy(h); return rc; };
  if (hipStreamCreate(&h->stream) != hipSuccess) { h->err = "hipStreamCreate failed"; return bail(FVWB_ERR_HIP); }
  const int nx0 = per_inst[0].nx;
  if (hipMalloc(&h->d_inst, sizeof(Inst) * ninst) != hipSuccess ||
      hipMalloc(&h->d_dev, sizeof(double) * 3 * h->cells) != hipSuccess ||
      hipMalloc(&h->d_ueqc, sizeof(double) * 3 * h->cells) != hipSuccess ||
      hipMalloc(&h->d_ueqf, sizeof(double) * 3 * h->faces) != hipSuccess ||
      hipMalloc(&h->d_reff, sizeof(double) * 3 * (nx0 + 1)) != hipSuccess ||
      hipMalloc(&h->d_stage, sizeof(double) * 3 * 3 * (size_t)(maxnx > nx0 ? maxnx : nx0)) != hipSuccess ||
      hipMalloc(&h->d_clk, sizeof(Clock) * ninst) != hipSuccess ||
      hipMalloc(&h->d_red, sizeof(unsigned long long)) != hipSuccess) {
    h->err = "device allocation failed";
    return bail(FVWB_ERR_OOM);
  }
  // face tables: the configured closed form (planes) and the reference's (AoS)
  std::vector<double> tf(3 * (size_t)h->faces), tr(3 * (size_t)(nx0 + 1));
  for (int b = 0; b < ninst; ++b) {
    const fvwb1d_params& p = per_inst[b];
    const Inst& I = h->inst[b];
    for (int f = 0; f <= I.nx; ++f) {
      double w[3], u[3];
      equilibrium_host(p, (double)f * I.dx, w);
      conservative(p.gamma, w, u);
      for (int v = 0; v < 3; ++v) tf[3 * I.toff + (size_t)v * (I.nx + 1) + f] = u[v];
    }
  }
  for (int f = 0; f <= nx0; ++f) {
    double w[3];
    equilibrium_ref(per_inst[0], (double)f * h->inst[0].dx, w);
    conservative(per_inst[0].gamma, w, &tr[3 * f]);
  }
  std::vector<Clock> ck(ninst, Clock{0.0, 0.0, 0, 0});
  if (hipMemcpy(h->d_inst, h->inst.data(), sizeof(Inst) * ninst, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->d_ueqf, tf.data(), sizeof(double) * tf.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->d_reff, tr.data(), sizeof(double) * tr.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->d_clk, ck.data(), sizeof(Clock) * ninst, hipMemcpyHostToDevice) != hipSuccess) {
    h->err = "copying tables failed";
    return bail(FVWB_ERR_HIP);
  }
  *out = h;
  return FVWB_OK;
}

extern "C" const char* fvwb1d_last_error(fvwb1d_handle h) { return h ? h->err.c_str() : g_err1d.c_str(); }

extern "C" int fvwb1d_compute_update(fvwb1d_handle h, const double* u, const double* w_eq, double* dudt) {
  if (!h || !u || !w_eq || !dudt) return fail1d(h, FVWB_ERR_ARG, "NULL argument");
  const fvwb1d_params& p = h->p[0];
  const int nx = p.nx;
  const size_t bytes = sizeof(double) * 3 * (size_t)nx;
  double *du = h->d_stage, *dw = du + 3 * (size_t)nx, *dd = dw + 3 * (size_t)nx;
  HIPC(h, hipMemcpyAsync(du, u, bytes, hipMemcpyHostToDevice, h->stream));
  HIPC(h, hipMemcpyAsync(dw, w_eq, bytes, hipMemcpyHostToDevice, h->stream));
  const double dx = p.boxlen / (double)nx;
  hipLaunchKernelGGL(fv1d_ref_update_kernel, dim3((nx + 255) / 256), dim3(256), 0, h->stream, du, dw, h->d_reff, dd,
                     nx, p.gamma, dx, 1 / dx, 1 / (double)nx);
  HIPC(h, hipGetLastError());
  HIPC(h, hipMemcpyAsync(dudt, dd, bytes, hipMemcpyDeviceToHost, h->stream));
  HIPC(h, hipStreamSynchronize(h->stream));
  return FVWB_OK;
}

extern "C" int fvwb1d_compute_update_sr(fvwb1d_handle h, const double* u, double* dudt) {
  if (!h || !u || !dudt) return fail1d(h, FVWB_ERR_ARG, "NULL argument");
  const fvwb1d_params& p = h->p[0];
  const int nx = p.nx;
  const size_t bytes = sizeof(double) * 3 * (size_t)nx;
  double *du = h->d_stage, *dd = du + 3 * (size_t)nx;
  HIPC(h, hipMemcpyAsync(du, u, bytes, hipMemcpyHostToDevice, h->stream));
  const double dx = p.boxlen / (double)nx;
  hipLaunchKernelGGL(fv1d_ref_update_sr_kernel, dim3((nx + 255) / 256), dim3(256), 0, h->stream, du, dd, nx,
                     p.gamma, dx, p.boxlen / dx, 1 / (double)nx);
  HIPC(h, hipGetLastError());
  HIPC(h, hipMemcpyAsync(dudt, dd, bytes, hipMemcpyDeviceToHost, h->stream));
  HIPC(h, hipStreamSynchronize(h->stream));
  return FVWB_OK;
}

static int plain_update(fvwb1d* h, bool program, int source, const double* u, double* dudt) {
  const fvwb1d_params& p = h->p[0];
  const int nx = p.nx;
  const size_t bytes = sizeof(double) * 3 * (size_t)nx;
  double *du = h->d_stage, *dd = du + 3 * (size_t)nx;
  HIPC(h, hipMemcpyAsync(du, u, bytes, hipMemcpyHostToDevice, h->stream));
  const double dx = p.boxlen / (double)nx;
  const dim3 grid((nx + 255) / 256), block(256);
  if (program)
    hipLaunchKernelGGL(fv1d_plain_update_kernel<true>, grid, block, 0, h->stream, du, dd, nx, p.gamma, dx, 1.0 / dx,
                       dx, p.bc == FVWB1D_BC_PERIODIC ? 1 : 0, source);
  else
    hipLaunchKernelGGL(fv1d_plain_update_kernel<false>, grid, block, 0, h->stream, du, dd, nx, p.gamma, dx, 1 / dx,
                       1 / (double)nx, 0, 2);
  HIPC(h, hipGetLastError());
  HIPC(h, hipMemcpyAsync(dudt, dd, bytes, hipMemcpyDeviceToHost, h->stream));
  HIPC(h, hipStreamSynchronize(h->stream));
  return FVWB_OK;
}

extern "C" int fvwb1d_compute_update_fvm(fvwb1d_handle h, const double* u, double* dudt) {
  if (!h || !u || !dudt) return fail1d(h, FVWB_ERR_ARG, "NULL argument");
  return plain_update(h, false, 2, u, dudt);
}

extern "C" int fvwb1d_fvm_compute_update(fvwb1d_handle h, int32_t source, const double* u, double* dudt) {
  if (!h || !u || !dudt) return fail1d(h, FVWB_ERR_ARG, "NULL argument");
  if (source != 1 && source != 2) return fail1d(h, FVWB_ERR_ARG, "source must be 1 (none) or 2 (gravity)");
  const int bc = h->p[0].bc;
  if (bc != FVWB1D_BC_PERIODIC && bc != FVWB1D_BC_ZERO_GRADIENT)
    return fail1d(h, FVWB_ERR_ARG, "fvm.f90's compute_update handles bc 1 (periodic) and 2 (zero gradient) only");
  return plain_update(h, true, source, u, dudt);
}

extern "C" int fvwb1d_compute_max_speed(fvwb1d_handle h, const double* u, double* cmax) {
  if (!h || !u || !cmax) return fail1d(h, FVWB_ERR_ARG, "NULL argument");
  const fvwb1d_params& p = h->p[0];
  const int nx = p.nx;
  HIPC(h, hipMemcpyAsync(h->d_stage, u, sizeof(double) * 3 * (size_t)nx, hipMemcpyHostToDevice, h->stream));
  HIPC(h, hipMemsetAsync(h->d_red, 0, sizeof(unsigned long long), h->stream));
  int blocks = (nx + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(fv1d_max_speed_kernel, dim3(blocks), dim3(256), 0, h->stream, h->d_stage, nx, p.gamma, h->d_red);
  HIPC(h, hipGetLastError());
  unsigned long long bits = 0;
  HIPC(h, hipMemcpyAsync(&bits, h->d_red, sizeof bits, hipMemcpyDeviceToHost, h->stream));
  HIPC(h, hipStreamSynchronize(h->stream));
  std::memcpy(cmax, &bits, sizeof bits);
  return FVWB_OK;
}

// Host-side conversion to the device layout of one instance: ueqc and dU planes.
static int put_instance(fvwb1d* h, int b, const double* u, const double* w_eq) {
  const fvwb1d_params& p = h->p[b];
  const Inst& I = h->inst[b];
  const int nx = I.nx;
  std::vector<double> ue(3 * (size_t)nx), d(3 * (size_t)nx);
  for (int i = 0; i < nx; ++i) {
    double c[3];
    conservative(p.gamma, &w_eq[3 * i], c);
    for (int v = 0; v < 3; ++v) {
      ue[(size_t)v * nx + i] = c[v];
      d[(size_t)v * nx + i] = u[3 * i + v] - c[v];
    }
  }
  HIPC(h, hipMemcpyAsync(h->d_ueqc + 3 * I.off, ue.data(), sizeof(double) * ue.size(), hipMemcpyHostToDevice, h->stream));
  HIPC(h, hipMemcpyAsync(h->d_dev + 3 * I.off, d.data(), sizeof(double) * d.size(), hipMemcpyHostToDevice, h->stream));
  Clock z{0.0, 0.0, 0, 0};
  HIPC(h, hipMemcpyAsync(h->d_clk + b, &z, sizeof z, hipMemcpyHostToDevice, h->stream));
  HIPC(h, hipStreamSynchronize(h->stream));
  h->ready[b] = 1;
  return FVWB_OK;
}

extern "C" int fvwb1d_upload(fvwb1d_handle h, int32_t inst, const double* u, const double* w_eq) {
  if (!h || !u || !w_eq) return fail1d(h, FVWB_ERR_ARG, "NULL argument");
  if (inst < 0 || inst >= h->ninst) return fail1d(h, FVWB_ERR_ARG, "instance index out of range");
  return put_instance(h, inst, u, w_eq);
}

extern "C" int fvwb1d_init_gaussian(fvwb1d_handle h, const double* amp, double x0, double width) {
  if (!h || !amp) return fail1d(h, FVWB_ERR_ARG, "NULL argument");
  if (!(width > 0.0)) return fail1d(h, FVWB_ERR_ARG, "width must be > 0");
  for (int b = 0; b < h->ninst; ++b) {
    const fvwb1d_params& p = h->p[b];
    const Inst& I = h->inst[b];
    std::vector<double> w(3 * (size_t)I.nx), u(3 * (size_t)I.nx);
    for (int i = 0; i < I.nx; ++i) {
      const double x = ((double)i + 0.5) * I.dx;
      equilibrium_host(p, x, &w[3 * i]);
      conservative(p.gamma, &w[3 * i], &u[3 * i]);
      u[3 * i + 2] += amp[b] * std::exp(-((x - x0) * (x - x0)) / (2.0 * width * width)) / (p.gamma - 1.0);
    }
    int rc = put_instance(h, b, u.data(), w.data());
    if (rc) return rc;
  }
  return FVWB_OK;
}

template <int RECON, int BC, int INTEG>
static void launch_run(fvwb1d* h, int C, int nsteps, double t_end) {
  const dim3 grid(h->ninst), block(BS);
  switch (C) {
#define FV1D_CASE(CC)                                                                                              \
  case CC:                                                                                                         \
    hipLaunchKernelGGL((fv1d_run_kernel<RECON, BC, INTEG, CC>), grid, block, h->lds_bytes, h->stream, h->d_inst,  \
                       h->d_dev, h->d_ueqc, h->d_ueqf, h->d_clk, nsteps, t_end);                                   \
    break;
    FV1D_CASE(1) FV1D_CASE(2) FV1D_CASE(4) FV1D_CASE(8)
#undef FV1D_CASE
  }
}

template <int RECON, int BC, int INTEG>
static hipError_t set_lds(size_t bytes) {
  hipError_t e = hipSuccess;
  const void* fns[] = {(const void*)fv1d_run_kernel<RECON, BC, INTEG, 1>, (const void*)fv1d_run_kernel<RECON, BC, INTEG, 2>,
                       (const void*)fv1d_run_kernel<RECON, BC, INTEG, 4>, (const void*)fv1d_run_kernel<RECON, BC, INTEG, 8>};
  for (const void* f : fns) {
    hipError_t r = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (r != hipSuccess) e = r;
  }
  return e;
}

template <int RECON, int BC>
static int dispatch_integ(fvwb1d* h, int nsteps, double t_end, bool big) {
  if (h->p[0].integrator == FVWB_SSPRK3) {
    if (big) HIPC(h, (set_lds<RECON, BC, FVWB_SSPRK3>(h->lds_bytes)));
    launch_run<RECON, BC, FVWB_SSPRK3>(h, h->cmax_per_lane, nsteps, t_end);
  } else {
    if (big) HIPC(h, (set_lds<RECON, BC, FVWB_RK2_HEUN>(h->lds_bytes)));
    launch_run<RECON, BC, FVWB_RK2_HEUN>(h, h->cmax_per_lane, nsteps, t_end);
  }
  return FVWB_OK;
}

extern "C" int fvwb1d_advance(fvwb1d_handle h, int32_t nsteps, double t_end) {
  if (!h) return fail1d(h, FVWB_ERR_ARG, "NULL handle");
  if (nsteps < 0 || std::isnan(t_end)) return fail1d(h, FVWB_ERR_ARG, "nsteps >= 0 and t_end not NaN required");
  for (int b = 0; b < h->ninst; ++b)
    if (!h->ready[b]) return fail1d(h, FVWB_ERR_STATE, "instance " + std::to_string(b) + " has no state");
  const fvwb1d_params& p = h->p[0];
  if (p.bc != FVWB1D_BC_WALL && p.bc != FVWB1D_BC_ZERO_GRADIENT)
    return fail1d(h, FVWB_ERR_ARG, "the configured scheme takes bc ZERO_GRADIENT or WALL");
  for (int b = 0; b < h->ninst; ++b)
    if (h->p[b].nx > MAX_NX) return fail1d(h, FVWB_ERR_ARG, "fvwb1d_advance supports nx <= 2048");
  int maxlds = 0;
  HIPC(h, hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, 0));
  if (h->lds_bytes > (size_t)maxlds)
    return fail1d(h, FVWB_ERR_ARG, "instance too large for one workgroup's LDS on this device");
  if (nsteps == 0) return FVWB_OK;
  const bool big = h->lds_bytes > 65536;
  int rc;
  const bool muscl = p.recon == FVWB_RECON_MUSCL_MINMOD, wall = p.bc == FVWB1D_BC_WALL;
  if (muscl && wall) rc = dispatch_integ<FVWB_RECON_MUSCL_MINMOD, FVWB1D_BC_WALL>(h, nsteps, t_end, big);
  else if (muscl) rc = dispatch_integ<FVWB_RECON_MUSCL_MINMOD, FVWB1D_BC_ZERO_GRADIENT>(h, nsteps, t_end, big);
  else if (wall) rc = dispatch_integ<FVWB_RECON_FIRST_ORDER, FVWB1D_BC_WALL>(h, nsteps, t_end, big);
  else rc = dispatch_integ<FVWB_RECON_FIRST_ORDER, FVWB1D_BC_ZERO_GRADIENT>(h, nsteps, t_end, big);
  if (rc) return rc;
  HIPC(h, hipGetLastError());
  std::vector<Clock> ck(h->ninst);
  HIPC(h, hipMemcpyAsync(ck.data(), h->d_clk, sizeof(Clock) * h->ninst, hipMemcpyDeviceToHost, h->stream));
  HIPC(h, hipStreamSynchronize(h->stream));
  for (int b = 0; b < h->ninst; ++b)
    if (ck[b].err)
      return fail1d(h, FVWB_ERR_NONPHYSICAL, "instance " + std::to_string(b) + ": rho<=0, p<=0, non-finite state or dt <= 0");
  return FVWB_OK;
}

extern "C" int fvwb1d_download(fvwb1d_handle h, int32_t inst, double* u) {
  if (!h || !u) return fail1d(h, FVWB_ERR_ARG, "NULL argument");
  if (inst < 0 || inst >= h->ninst) return fail1d(h, FVWB_ERR_ARG, "instance index out of range");
  if (!h->ready[inst]) return fail1d(h, FVWB_ERR_STATE, "instance has no state");
  const Inst& I = h->inst[inst];
  const int nx = I.nx;
  double* tmp = h->d_stage;
  hipLaunchKernelGGL(fv1d_to_aos_kernel, dim3((nx + 255) / 256), dim3(256), 0, h->stream, h->d_dev + 3 * I.off,
                     h->d_ueqc + 3 * I.off, tmp, nx);
  HIPC(h, hipGetLastError());
  HIPC(h, hipMemcpyAsync(u, tmp, sizeof(double) * 3 * (size_t)nx, hipMemcpyDeviceToHost, h->stream));
  HIPC(h, hipStreamSynchronize(h->stream));
  return FVWB_OK;
}

extern "C" int fvwb1d_get_time(fvwb1d_handle h, int32_t inst, double* t, double* dt_last, int32_t* steps_done) {
  if (!h) return fail1d(h, FVWB_ERR_ARG, "NULL handle");
  if (inst < 0 || inst >= h->ninst) return fail1d(h, FVWB_ERR_ARG, "instance index out of range");
  Clock c{};
  HIPC(h, hipMemcpyAsync(&c, h->d_clk + inst, sizeof c, hipMemcpyDeviceToHost, h->stream));
  HIPC(h, hipStreamSynchronize(h->stream));
  if (t) *t = c.t;
  if (dt_last) *dt_last = c.dt_last;
  if (steps_done) *steps_done = c.steps;
  return FVWB_OK;
}

// ---- checkpoint / resume: every instance's deviation, centre equilibrium and clock
extern "C" int fvwb1d_checkpoint_save(fvwb1d_handle h, const char* path) {
  if (!h || !path) return fail1d(h, FVWB_ERR_ARG, "NULL argument");
  for (int b = 0; b < h->ninst; ++b)
    if (!h->ready[b]) return fail1d(h, FVWB_ERR_STATE, "checkpoint before every instance was uploaded");
  const fvwb_ckpt::Segment seg[3] = {{h->d_dev, sizeof(double) * 3 * (size_t)h->cells},
                                     {h->d_ueqc, sizeof(double) * 3 * (size_t)h->cells},
                                     {h->d_clk, sizeof(Clock) * (size_t)h->ninst}};
  std::string err;
  const int rc = fvwb_ckpt::save(path, fvwb_ckpt::FV1D, h->p.data(), sizeof(fvwb1d_params) * h->p.size(), 0, seg, 3,
                                 h->stream, err);
  if (rc) h->err = err;
  return rc;
}

extern "C" int fvwb1d_checkpoint_load(fvwb1d_handle h, const char* path) {
  if (!h || !path) return fail1d(h, FVWB_ERR_ARG, "NULL argument");
  const fvwb_ckpt::Segment seg[3] = {{h->d_dev, sizeof(double) * 3 * (size_t)h->cells},
                                     {h->d_ueqc, sizeof(double) * 3 * (size_t)h->cells},
                                     {h->d_clk, sizeof(Clock) * (size_t)h->ninst}};
  std::string err;
  int32_t parity = 0;
  const int rc = fvwb_ckpt::load(path, fvwb_ckpt::FV1D, h->p.data(), sizeof(fvwb1d_params) * h->p.size(), &parity,
                                 seg, 3, h->stream, err);
  for (int b = 0; b < h->ninst; ++b) h->ready[b] = rc == FVWB_OK;
  if (rc) h->err = err;
  return rc;
}

extern "C" int fvwb1d_get_stream(fvwb1d_handle h, void** stream) {
  if (!h || !stream) return fail1d(h, FVWB_ERR_ARG, "NULL argument");
  *stream = (void*)h->stream;
  return FVWB_OK;
}

extern "C" int fvwb1d_compute_residual(fvwb1d_handle h, int32_t inst, const double* u, const double* w_eq,
                                       double* dudt) {
  if (!h || !u || !w_eq || !dudt) return fail1d(h, FVWB_ERR_ARG, "NULL argument");
  if (inst < 0 || inst >= h->ninst) return fail1d(h, FVWB_ERR_ARG, "instance index out of range");
  const fvwb1d_params& p = h->p[inst];
  if (p.bc != FVWB1D_BC_WALL && p.bc != FVWB1D_BC_ZERO_GRADIENT)
    return fail1d(h, FVWB_ERR_ARG, "the configured scheme takes bc ZERO_GRADIENT or WALL");
  const Inst& I = h->inst[inst];
  const int nx = I.nx;
  std::vector<double> d(3 * (size_t)nx);
  for (int i = 0; i < nx; ++i) {
    double c[3];
    conservative(p.gamma, &w_eq[3 * i], c);
    for (int v = 0; v < 3; ++v) d[(size_t)v * nx + i] = u[3 * i + v] - c[v];
  }
  double *dd = h->d_stage, *dl = dd + 3 * (size_t)nx;
  HIPC(h, hipMemcpyAsync(dd, d.data(), sizeof(double) * d.size(), hipMemcpyHostToDevice, h->stream));
  const dim3 grid((nx + 255) / 256), block(256);
  const double* tab = h->d_ueqf + 3 * I.toff;
  const bool muscl = p.recon == FVWB_RECON_MUSCL_MINMOD, wall = p.bc == FVWB1D_BC_WALL;
  if (muscl && wall)
    hipLaunchKernelGGL((fv1d_residual_kernel<FVWB_RECON_MUSCL_MINMOD, FVWB1D_BC_WALL>), grid, block, 0, h->stream, dd, tab, dl, nx, p.gamma, I.odx, p.g);
  else if (muscl)
    hipLaunchKernelGGL((fv1d_residual_kernel<FVWB_RECON_MUSCL_MINMOD, FVWB1D_BC_ZERO_GRADIENT>), grid, block, 0, h->stream, dd, tab, dl, nx, p.gamma, I.odx, p.g);
  else if (wall)
    hipLaunchKernelGGL((fv1d_residual_kernel<FVWB_RECON_FIRST_ORDER, FVWB1D_BC_WALL>), grid, block, 0, h->stream, dd, tab, dl, nx, p.gamma, I.odx, p.g);
  else
    hipLaunchKernelGGL((fv1d_residual_kernel<FVWB_RECON_FIRST_ORDER, FVWB1D_BC_ZERO_GRADIENT>), grid, block, 0, h->stream, dd, tab, dl, nx, p.gamma, I.odx, p.g);
  HIPC(h, hipGetLastError());
  HIPC(h, hipMemcpyAsync(dudt, dl, sizeof(double) * 3 * (size_t)nx, hipMemcpyDeviceToHost, h->stream));
  HIPC(h, hipStreamSynchronize(h->stream));
  return FVWB_OK;
}
