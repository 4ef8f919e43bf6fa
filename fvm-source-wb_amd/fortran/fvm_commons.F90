! fvm_commons — the set-up module of the stand-alone 1D program fvm.f90 for the
! driver fvm_fvwb and its drop-ins (names, kinds and default values of
! fvm_commons.f90:1-21).  nx, ninit, bc and source can be overridden at build time
! (-DFVWB_NX=..., -DFVWB_NINIT=..., -DFVWB_BC=..., -DFVWB_SOURCE=...).
#ifndef FVWB_NX
#define FVWB_NX 200
#endif
#ifndef FVWB_NINIT
#define FVWB_NINIT 4
#endif
#ifndef FVWB_BC
#define FVWB_BC 2
#endif
#ifndef FVWB_SOURCE
#define FVWB_SOURCE 2
#endif
module fvm_commons
  implicit none
  integer, parameter :: n = 3
  character(len=3), parameter :: integrator = 'RK4'
  integer, parameter :: nx = FVWB_NX
  integer, parameter :: nvar = 3
  integer, parameter :: riemann = 2
  logical, parameter :: use_limiter = .true.
  integer, parameter :: ninit = FVWB_NINIT   ! 4: Sod
  integer, parameter :: bc = FVWB_BC         ! 1 periodic, 2 zero gradient
  integer, parameter :: source = FVWB_SOURCE ! 1 none, 2 gravity phi = x
  ! default-real literals on purpose, as in the reference set-up
  real(kind=8) :: tend = 0.2
  real(kind=8) :: boxlen = 1
  real(kind=8) :: gamma = 1.4
end module fvm_commons
