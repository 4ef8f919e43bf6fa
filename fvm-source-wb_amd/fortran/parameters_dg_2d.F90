! parameters_dg_2d — the set-up module of the reference's 2D DG program and of
! the drop-ins that replace its procedures (names, kinds and default values of
! 2d/parameters_dg_2d.f90, so code written against that module compiles
! unchanged against this one).
!
! The compile-time switches there are constants here too; each can be
! overridden at build time so one source builds the shipped set-up and its
! variants: -DFVWB_N=... (nx = ny), -DFVWB_M=... (mx = my),
! -DFVWB_FLUX_TYPE="'hll2'", -DFVWB_SOLVER="'SS4'", -DFVWB_SOURCE=...,
! -DFVWB_GRAD_PHI=..., -DFVWB_BC=..., -DFVWB_NINIT=..., -DFVWB_NEQ=....
! tend, the box, gamma, cfl, eps and eta stay variables, as there.
#ifndef FVWB_N
#define FVWB_N 8
#endif
#ifndef FVWB_M
#define FVWB_M 2
#endif
#ifndef FVWB_FLUX_TYPE
#define FVWB_FLUX_TYPE 'llf'
#endif
#ifndef FVWB_SOLVER
#define FVWB_SOLVER 'RK4'
#endif
#ifndef FVWB_SOURCE
#define FVWB_SOURCE 1
#endif
#ifndef FVWB_GRAD_PHI
#define FVWB_GRAD_PHI 2
#endif
#ifndef FVWB_BC
#define FVWB_BC 1
#endif
#ifndef FVWB_NINIT
#define FVWB_NINIT 1
#endif
#ifndef FVWB_NEQ
#define FVWB_NEQ 3
#endif
module parameters_dg_2d
  implicit none
  integer, parameter :: nx = FVWB_N, ny = FVWB_N
  integer, parameter :: mx = FVWB_M, my = FVWB_M
  integer, parameter :: gll = (2 * (mx - 1) + 3) / 2
  integer, parameter :: nvar = 4
  integer, parameter :: riemann = 2
  logical, parameter :: use_limiter = .true.
  logical, parameter :: make_movie = .true.
  character(len=3), parameter :: solver = FVWB_SOLVER          ! 'RK4', 'SS4', 'EQL' or 'DEB'
  character(len=3), parameter :: limiter_type = 'ONP'          ! compute_positivity
  character(len=4), parameter :: flux_type = FVWB_FLUX_TYPE    ! 'llf1', 'hll2', 'hllc'; others: none
  integer, parameter :: ninit = FVWB_NINIT
  integer, parameter :: bc = FVWB_BC
  integer, parameter :: nequilibrium = FVWB_NEQ
  integer, parameter :: source = FVWB_SOURCE
  integer, parameter :: grad_phi_case = FVWB_GRAD_PHI
  ! default-real literals on purpose, as the reference writes them: the
  ! drop-ins then compute with the same rounded constants
  real(kind=8) :: tend = 1.0
  real(kind=8) :: boxlen_x = 1.0
  real(kind=8) :: boxlen_y = 1.0
  real(kind=8) :: gamma = 1.4
  real(kind=8) :: cfl = 0.2
  real(kind=8) :: eps = 1e-10
  real(kind=8) :: eta = 0.1
  real(kind=8), dimension(1:mx) :: x_quad, w_x_quad
  real(kind=8), dimension(1:mx) :: y_quad, w_y_quad
  real(kind=8), dimension(1:gll) :: x_gll_quad, w_gll_quad
  real(kind=8) :: M = 0
  integer, parameter :: interval = 50
end module parameters_dg_2d
