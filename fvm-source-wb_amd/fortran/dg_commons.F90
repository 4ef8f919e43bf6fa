! dg_commons — the set-up module of the 1D DG program (dg_with_source.f90) for
! the driver dg1d_fvwb and the drop-in compute_update_exact_delta (names, kinds
! and default values of dg_commons.f90:1-25).  n, nx, riemann and ninit can be
! overridden at build time (-DFVWB_N=..., -DFVWB_NX=..., -DFVWB_RIEMANN=...,
! -DFVWB_NINIT=...).
#ifndef FVWB_N
#define FVWB_N 3
#endif
#ifndef FVWB_NX
#define FVWB_NX 128
#endif
#ifndef FVWB_RIEMANN
#define FVWB_RIEMANN 2
#endif
#ifndef FVWB_NINIT
#define FVWB_NINIT 8
#endif
module dg_commons
  implicit none
  integer, parameter :: n = FVWB_N                 ! modes = Gauss points
  character(len=3), parameter :: integrator = 'RKi'
  integer, parameter :: nx = FVWB_NX
  integer, parameter :: nquad = n
  integer, parameter :: nvar = 3
  integer, parameter :: riemann = FVWB_RIEMANN     ! 1 LLF, 2 HLLC
  logical, parameter :: use_limiter = .false.
  integer, parameter :: ninit = FVWB_NINIT         ! 7 rest, 8 rest plus a pulse of size pert
  integer, parameter :: bc = 5
  integer, parameter :: source = 2                 ! 1 none, 2 gravity phi = x
  ! default-real literals on purpose, as in the reference set-up
  real(kind=8) :: tend = 0.2
  real(kind=8) :: boxlen = 1.0
  real(kind=8) :: gamma = 1.4
  real(kind=8) :: pert = 1e-8
  real(kind=8), dimension(1:nquad) :: chsi_quad, w_quad
  character(len=7) :: finaloutput = 'dg31e-8'
end module dg_commons
