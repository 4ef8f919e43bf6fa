! parameters — the set-up module of the 1D benchmark driver and drop-ins
! (names, kinds and default values of parameters.f90:1-22).  nx, ninit,
! nequilibrium and make_movie can be overridden at build time (-DFVWB_NX=...,
! -DFVWB_NINIT=..., -DFVWB_NEQUILIBRIUM=..., -DFVWB_MAKE_MOVIE=.true.).
#ifndef FVWB_NX
#define FVWB_NX 128
#endif
#ifndef FVWB_NINIT
#define FVWB_NINIT 2
#endif
#ifndef FVWB_NEQUILIBRIUM
#define FVWB_NEQUILIBRIUM 2
#endif
#ifndef FVWB_MAKE_MOVIE
#define FVWB_MAKE_MOVIE .false.
#endif
module parameters
  implicit none
  integer, parameter :: nx = FVWB_NX
  integer, parameter :: nvar = 3
  integer, parameter :: riemann = 2
  logical, parameter :: use_limiter = .true.
  logical, parameter :: make_movie = FVWB_MAKE_MOVIE
  character(len=3), parameter :: solver = 'EQL'     ! the exact deviation path the drop-ins replace
  integer, parameter :: ninit = FVWB_NINIT          ! 1: exp(-x); 2: 1 + pulse of amplitude eta; 3: polytropic
  integer, parameter :: bc = 2
  integer, parameter :: nequilibrium = FVWB_NEQUILIBRIUM
  ! default-real literals on purpose, as in the reference set-up
  real(kind=8) :: tend = 0.20
  real(kind=8) :: boxlen = 1.0
  real(kind=8) :: gamma = 1.4
  real(kind=8) :: eta = 1e-8
  character(len=7) :: finaloutput = 'wb11e-8'
end module parameters
