! parameters_rd — the module benchmark_rd.f90 uses (:5) and the reference does
! not ship: the names its routines read (nx, ny, mx, my, nvar, boxlen_x,
! boxlen_y, gamma, bc, cfl, tend, ninit, solver, use_limiter, make_movie, eta),
! with defaults in the style of parameters_2d.f90 (default-real literals, as
! there).  nx, ny, ninit and bc can be overridden at build time
! (-DFVWB_NX=..., -DFVWB_NY=..., -DFVWB_NINIT=..., -DFVWB_BC=...).
#ifndef FVWB_NX
#define FVWB_NX 65
#endif
#ifndef FVWB_NY
#define FVWB_NY 65
#endif
#ifndef FVWB_NINIT
#define FVWB_NINIT 8
#endif
#ifndef FVWB_BC
#define FVWB_BC 3
#endif
module parameters_rd
  implicit none
  integer, parameter :: nx = FVWB_NX
  integer, parameter :: ny = FVWB_NY
  integer, parameter :: mx = 1
  integer, parameter :: my = 1
  integer, parameter :: nvar = 4
  integer, parameter :: ninit = FVWB_NINIT      ! 1 density bump, 2 rest state, 8 rest state + pressure bump
  integer, parameter :: nequilibrium = 2
  integer, parameter :: bc = FVWB_BC            ! 2 periodic copies, 3 transmissive copies
  character(len=3), parameter :: solver = 'RD1'
  logical, parameter :: use_limiter = .false.   ! lim is commented out in compute_update_rd
  logical, parameter :: make_movie = .false.
  real(kind=8) :: boxlen_x = 1
  real(kind=8) :: boxlen_y = 1
  real(kind=8) :: gamma = 1.4
  real(kind=8) :: cfl = 0.5
  real(kind=8) :: tend = 0.1
  real(kind=8) :: eta = 0.01
end module parameters_rd
