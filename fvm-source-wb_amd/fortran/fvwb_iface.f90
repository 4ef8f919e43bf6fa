! fvwb_iface — Fortran 2008 iso_c_binding interfaces to libfvwb (include/fvwb.h).
!
! Every procedure below binds one extern "C" entry point of the C ABI with the
! same name and argument order; arrays are passed as assumed-size real(c_double)
! in the Fortran layout u(1:nvar,1:nx,1:ny), which is exactly the C layout the
! library documents (variable fastest, then x, then y).  Status codes are the
! FVWB_* constants; fvwb_check turns a nonzero status into `error stop` with the
! library's message, so a caller written like the reference (no status
! handling) still stops loudly on a device error.
module fvwb_iface
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: FVWB_OK = 0, FVWB_ERR_ARG = -1, FVWB_ERR_HIP = -2, &
       FVWB_ERR_OOM = -3, FVWB_ERR_NONPHYSICAL = -4, FVWB_ERR_STATE = -5
  integer(c_int32_t), parameter, public :: FVWB_BC_FROZEN_RING = 0, FVWB_BC_REFLECT = 1
  integer(c_int32_t), parameter, public :: FVWB_RECON_FIRST_ORDER = 0, FVWB_RECON_MUSCL_MINMOD = 1
  integer(c_int32_t), parameter, public :: FVWB_FLUX_RUSANOV = 0, FVWB_FLUX_HLL = 1
  integer(c_int32_t), parameter, public :: FVWB_RK2_HEUN = 0, FVWB_SSPRK3 = 1
  integer(c_int32_t), parameter, public :: FVWB_FP64 = 0, FVWB_FP32 = 1

  ! struct fvwb_params: eight int32 then seven doubles (no padding: 32 + 56 bytes)
  type, bind(C), public :: fvwb_params
     integer(c_int32_t) :: nvar, nx, ny, bc, recon, flux, integrator, precision
     real(c_double) :: gamma, boxlen_x, boxlen_y, cfl, rho0, p0, g
  end type fvwb_params

  public :: fvwb_params_reference_2d, fvwb_create, fvwb_destroy, fvwb_last_error
  public :: fvwb_compute_update_exact, fvwb_compute_update, fvwb_compute_max_speed
  public :: fvwb_upload, fvwb_init_gaussian, fvwb_advance, fvwb_download, fvwb_download_deviation
  public :: fvwb_get_time
  public :: fvwb_batch_create, fvwb_batch_destroy, fvwb_batch_last_error, fvwb_batch_init_gaussian
  public :: fvwb_batch_upload, fvwb_batch_reset_time, fvwb_batch_advance, fvwb_batch_download
  public :: fvwb_batch_download_deviation, fvwb_batch_get_time
  public :: fvwb_check, fvwb_batch_check, fvwb_message
  public :: fvwb_checkpoint_save, fvwb_checkpoint_load, fvwb1d_checkpoint_save, fvwb1d_checkpoint_load
  public :: fvwb_dg_checkpoint_save, fvwb_dg_checkpoint_load, fvwb_dg1d_checkpoint_save, fvwb_dg1d_checkpoint_load

  ! ---- 1D (benchmark_1d.f90) ----
  integer(c_int32_t), parameter, public :: FVWB1D_BC_PERIODIC = 1, FVWB1D_BC_ZERO_GRADIENT = 2, &
       FVWB1D_BC_REFLEXIVE = 3, FVWB1D_BC_WALL = 4

  ! ---- 1D modal DG (dg_with_source.f90) ----
  integer(c_int32_t), parameter, public :: FVWB_DG1D_RIEMANN_LLF = 1, FVWB_DG1D_RIEMANN_HLLC = 2

  ! struct fvwb_dg1d_params: six int32 then two doubles (24 + 16 bytes)
  type, bind(C), public :: fvwb_dg1d_params
     integer(c_int32_t) :: nvar, nx, n, riemann, source, pad
     real(c_double) :: gamma, boxlen
  end type fvwb_dg1d_params

  public :: fvwb_dg1d_params_reference, fvwb_dg1d_create, fvwb_dg1d_destroy, fvwb_dg1d_last_error
  public :: fvwb_dg1d_compute_update_exact_delta, fvwb_dg1d_upload, fvwb_dg1d_advance
  public :: fvwb_dg1d_download, fvwb_dg1d_get_time, fvwb_dg1d_check

  ! ---- 2D residual distribution (benchmark_rd.f90) ----
  integer(c_int32_t), parameter, public :: FVWB_RD_BC_PERIODIC = 2, FVWB_RD_BC_TRANSMISSIVE = 3

  ! struct fvwb_rd_params: four int32 then four doubles (16 + 32 bytes)
  type, bind(C), public :: fvwb_rd_params
     integer(c_int32_t) :: nx, ny, bc, pad
     real(c_double) :: gamma, boxlen_x, boxlen_y, cfl
  end type fvwb_rd_params

  public :: fvwb_rd_create, fvwb_rd_destroy, fvwb_rd_last_error, fvwb_rd_compute_update
  public :: fvwb_rd_compute_max_speed, fvwb_rd_upload, fvwb_rd_advance, fvwb_rd_download
  public :: fvwb_rd_get_time, fvwb_rd_check, fvwb_rd_checkpoint_save, fvwb_rd_checkpoint_load

  ! ---- 2D modal DG in deviation form (the configured P2 scheme, FP32 device path) ----
  integer(c_int32_t), parameter, public :: FVWB_FLUX_HLLC = 2
  integer(c_int32_t), parameter, public :: FVWB_DG_LIMITER_NONE = 0, FVWB_DG_LIMITER_TVB = 1, &
       FVWB_DG_LIMITER_POSITIVITY = 2, FVWB_DG_LIMITER_TVB_POSITIVITY = 3, FVWB_DG_LIMITER_CHARACTERISTIC = 4, &
       FVWB_DG_LIMITER_CHARACTERISTIC_POSITIVITY = 5
  integer(c_int32_t), parameter, public :: FVWB_DG_SSPRK3 = 0, FVWB_DG_SSPRK54 = 1

  ! struct fvwb_dg_params: eight int32 then eight doubles (32 + 64 bytes)
  type, bind(C), public :: fvwb_dg_params
     integer(c_int32_t) :: nvar, nx, ny, degree, flux, limiter, bc, integrator
     real(c_double) :: gamma, boxlen_x, boxlen_y, cfl, rho0, p0, g, tvb_m
  end type fvwb_dg_params

  public :: fvwb_dg_create, fvwb_dg_destroy, fvwb_dg_last_error, fvwb_dg_compute_residual, fvwb_dg_upload
  public :: fvwb_dg_init_fourier, fvwb_dg_advance, fvwb_dg_download, fvwb_dg_get_time, fvwb_dg_check

  ! ---- the reference's own 2D DG program (2d/benchmark_2d_dg.f90, FP64) ----
  integer(c_int32_t), parameter, public :: FVWB_DGQ_FLUX_NONE = 0, FVWB_DGQ_FLUX_LLF1 = 1, &
       FVWB_DGQ_FLUX_HLL2 = 2, FVWB_DGQ_FLUX_HLLC = 3
  integer(c_int32_t), parameter, public :: FVWB_DGQ_SOLVER_RK4 = 0, FVWB_DGQ_SOLVER_SS4 = 1, &
       FVWB_DGQ_SOLVER_EQL = 2, FVWB_DGQ_SOLVER_DEB = 3

  ! struct fvwb_dgq_params: ten int32 then five doubles (40 + 40 bytes)
  type, bind(C), public :: fvwb_dgq_params
     integer(c_int32_t) :: nx, ny, order, flux_type, source, grad_phi_case, bc, ninit, solver, use_limiter
     real(c_double) :: gamma, boxlen_x, boxlen_y, cfl, eps
  end type fvwb_dgq_params

  public :: fvwb_dgq_params_reference, fvwb_dgq_gl_quadrature, fvwb_dgq_create, fvwb_dgq_destroy
  public :: fvwb_dgq_last_error, fvwb_dgq_compute_update, fvwb_dgq_compute_max_speed, fvwb_dgq_apply_limiter
  public :: fvwb_dgq_modes_from_nodes, fvwb_dgq_nodes_from_modes, fvwb_dgq_upload, fvwb_dgq_advance
  public :: fvwb_dgq_download, fvwb_dgq_get_time, fvwb_dgq_compute_error, fvwb_dgq_check

  ! struct fvwb1d_params: six int32 then four doubles (24 + 32 bytes)
  type, bind(C), public :: fvwb1d_params
     integer(c_int32_t) :: nvar, nx, bc, recon, integrator, nequilibrium
     real(c_double) :: gamma, boxlen, cfl, g
  end type fvwb1d_params

  public :: fvwb1d_params_reference, fvwb1d_create, fvwb1d_destroy, fvwb1d_last_error
  public :: fvwb1d_compute_update, fvwb1d_compute_update_sr, fvwb1d_compute_update_fvm
  public :: fvwb1d_fvm_compute_update, fvwb1d_compute_max_speed
  public :: fvwb1d_upload, fvwb1d_init_gaussian
  public :: fvwb1d_advance, fvwb1d_download, fvwb1d_get_time, fvwb1d_check

  interface
     integer(c_int) function fvwb_params_reference_2d(p, nx, ny, nequilibrium) &
          bind(C, name='fvwb_params_reference_2d')
       import :: c_int, c_int32_t, fvwb_params
       type(fvwb_params), intent(out) :: p
       integer(c_int32_t), value :: nx, ny, nequilibrium
     end function fvwb_params_reference_2d

     integer(c_int) function fvwb_create(h, p) bind(C, name='fvwb_create')
       import :: c_int, c_ptr, fvwb_params
       type(c_ptr), intent(out) :: h
       type(fvwb_params), intent(in) :: p
     end function fvwb_create

     integer(c_int) function fvwb_destroy(h) bind(C, name='fvwb_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function fvwb_destroy

     type(c_ptr) function fvwb_last_error(h) bind(C, name='fvwb_last_error')
       import :: c_ptr
       type(c_ptr), value :: h
     end function fvwb_last_error

     ! drop-in for benchmark_2d.f90:465-618 (compute_update_exact(u,w_eq,dudt))
     integer(c_int) function fvwb_compute_update_exact(h, u, w_eq, dudt) &
          bind(C, name='fvwb_compute_update_exact')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: u(*), w_eq(*)
       real(c_double), intent(out) :: dudt(*)
     end function fvwb_compute_update_exact

     ! drop-in for benchmark_2d.f90:370-463 (compute_update(u,w_eq,dudt), plain, non-WB)
     integer(c_int) function fvwb_compute_update(h, u, w_eq, dudt) bind(C, name='fvwb_compute_update')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: u(*), w_eq(*)
       real(c_double), intent(out) :: dudt(*)
     end function fvwb_compute_update

     ! drop-in for benchmark_2d.f90:264-279 (compute_max_speed(u,cmax))
     integer(c_int) function fvwb_compute_max_speed(h, u, cmax) bind(C, name='fvwb_compute_max_speed')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: u(*)
       real(c_double), intent(out) :: cmax
     end function fvwb_compute_max_speed

     integer(c_int) function fvwb_upload(h, u, w_eq) bind(C, name='fvwb_upload')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: u(*), w_eq(*)
     end function fvwb_upload

     integer(c_int) function fvwb_init_gaussian(h, amp, x0, y0, width) bind(C, name='fvwb_init_gaussian')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), value :: amp, x0, y0, width
     end function fvwb_init_gaussian

     integer(c_int) function fvwb_advance(h, nsteps, t_end, t_out, steps_done) bind(C, name='fvwb_advance')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: nsteps
       real(c_double), value :: t_end
       real(c_double), intent(out) :: t_out
       integer(c_int32_t), intent(out) :: steps_done
     end function fvwb_advance

     integer(c_int) function fvwb_download(h, u) bind(C, name='fvwb_download')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: u(*)
     end function fvwb_download

     integer(c_int) function fvwb_download_deviation(h, du) bind(C, name='fvwb_download_deviation')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: du(*)
     end function fvwb_download_deviation

     integer(c_int) function fvwb_get_time(h, t, dt_last, steps_done) bind(C, name='fvwb_get_time')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: t, dt_last
       integer(c_int32_t), intent(out) :: steps_done
     end function fvwb_get_time

     integer(c_int) function fvwb_batch_create(h, ninst, per_inst) bind(C, name='fvwb_batch_create')
       import :: c_int, c_int32_t, c_ptr, fvwb_params
       type(c_ptr), intent(out) :: h
       integer(c_int32_t), value :: ninst
       type(fvwb_params), intent(in) :: per_inst(*)
     end function fvwb_batch_create

     integer(c_int) function fvwb_batch_destroy(h) bind(C, name='fvwb_batch_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function fvwb_batch_destroy

     type(c_ptr) function fvwb_batch_last_error(h) bind(C, name='fvwb_batch_last_error')
       import :: c_ptr
       type(c_ptr), value :: h
     end function fvwb_batch_last_error

     integer(c_int) function fvwb_batch_init_gaussian(h, amp, x0, y0, width) &
          bind(C, name='fvwb_batch_init_gaussian')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: amp(*)
       real(c_double), value :: x0, y0, width
     end function fvwb_batch_init_gaussian

     integer(c_int) function fvwb_batch_upload(h, inst, u, w_eq) bind(C, name='fvwb_batch_upload')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: inst
       real(c_double), intent(in) :: u(*), w_eq(*)
     end function fvwb_batch_upload

     integer(c_int) function fvwb_batch_reset_time(h) bind(C, name='fvwb_batch_reset_time')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function fvwb_batch_reset_time

     integer(c_int) function fvwb_batch_advance(h, nsteps, t_end) bind(C, name='fvwb_batch_advance')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: nsteps
       real(c_double), value :: t_end
     end function fvwb_batch_advance

     integer(c_int) function fvwb_batch_download(h, inst, u) bind(C, name='fvwb_batch_download')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: inst
       real(c_double), intent(out) :: u(*)
     end function fvwb_batch_download

     integer(c_int) function fvwb_batch_download_deviation(h, inst, du) &
          bind(C, name='fvwb_batch_download_deviation')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: inst
       real(c_double), intent(out) :: du(*)
     end function fvwb_batch_download_deviation

     integer(c_int) function fvwb_batch_get_time(h, inst, t, dt_last, steps_done) &
          bind(C, name='fvwb_batch_get_time')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: inst
       real(c_double), intent(out) :: t, dt_last
       integer(c_int32_t), intent(out) :: steps_done
     end function fvwb_batch_get_time

     integer(c_int) function fvwb1d_params_reference(p, nx, nequilibrium) bind(C, name='fvwb1d_params_reference')
       import :: c_int, c_int32_t, fvwb1d_params
       type(fvwb1d_params), intent(out) :: p
       integer(c_int32_t), value :: nx, nequilibrium
     end function fvwb1d_params_reference

     integer(c_int) function fvwb1d_create(h, ninst, per_inst) bind(C, name='fvwb1d_create')
       import :: c_int, c_int32_t, c_ptr, fvwb1d_params
       type(c_ptr), intent(out) :: h
       integer(c_int32_t), value :: ninst
       type(fvwb1d_params), intent(in) :: per_inst(*)
     end function fvwb1d_create

     integer(c_int) function fvwb1d_destroy(h) bind(C, name='fvwb1d_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function fvwb1d_destroy

     type(c_ptr) function fvwb1d_last_error(h) bind(C, name='fvwb1d_last_error')
       import :: c_ptr
       type(c_ptr), value :: h
     end function fvwb1d_last_error

     ! drop-in for benchmark_1d.f90:263-377 (compute_update(u,w_eq,dudt), solver 'EQL')
     integer(c_int) function fvwb1d_compute_update(h, u, w_eq, dudt) bind(C, name='fvwb1d_compute_update')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: u(*), w_eq(*)
       real(c_double), intent(out) :: dudt(*)
     end function fvwb1d_compute_update

     ! drop-in for benchmark_1d.f90:553-747 (compute_update_sr(u,w_eq,dudt), solver 'WB1')
     integer(c_int) function fvwb1d_compute_update_sr(h, u, dudt) bind(C, name='fvwb1d_compute_update_sr')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: u(*)
       real(c_double), intent(out) :: dudt(*)
     end function fvwb1d_compute_update_sr

     ! drop-in for benchmark_1d.f90:454-549 (compute_update_fvm(u,w_eq,dudt), solver 'FVM')
     integer(c_int) function fvwb1d_compute_update_fvm(h, u, dudt) bind(C, name='fvwb1d_compute_update_fvm')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: u(*)
       real(c_double), intent(out) :: dudt(*)
     end function fvwb1d_compute_update_fvm

     ! drop-in for fvm.f90:188-251 (compute_update(u,dudt) of the stand-alone program)
     integer(c_int) function fvwb1d_fvm_compute_update(h, source, u, dudt) &
          bind(C, name='fvwb1d_fvm_compute_update')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: source
       real(c_double), intent(in) :: u(*)
       real(c_double), intent(out) :: dudt(*)
     end function fvwb1d_fvm_compute_update

     ! drop-in for benchmark_1d.f90:157-169 (compute_max_speed(u,cmax))
     integer(c_int) function fvwb1d_compute_max_speed(h, u, cmax) bind(C, name='fvwb1d_compute_max_speed')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: u(*)
       real(c_double), intent(out) :: cmax
     end function fvwb1d_compute_max_speed

     integer(c_int) function fvwb1d_upload(h, inst, u, w_eq) bind(C, name='fvwb1d_upload')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: inst
       real(c_double), intent(in) :: u(*), w_eq(*)
     end function fvwb1d_upload

     integer(c_int) function fvwb1d_init_gaussian(h, amp, x0, width) bind(C, name='fvwb1d_init_gaussian')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: amp(*)
       real(c_double), value :: x0, width
     end function fvwb1d_init_gaussian

     integer(c_int) function fvwb1d_advance(h, nsteps, t_end) bind(C, name='fvwb1d_advance')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: nsteps
       real(c_double), value :: t_end
     end function fvwb1d_advance

     integer(c_int) function fvwb1d_download(h, inst, u) bind(C, name='fvwb1d_download')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: inst
       real(c_double), intent(out) :: u(*)
     end function fvwb1d_download

     integer(c_int) function fvwb1d_get_time(h, inst, t, dt_last, steps_done) bind(C, name='fvwb1d_get_time')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: inst
       real(c_double), intent(out) :: t, dt_last
       integer(c_int32_t), intent(out) :: steps_done
     end function fvwb1d_get_time

     integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
       import :: c_size_t, c_ptr
       type(c_ptr), value :: s
     end function c_strlen
     ! ---- checkpoint / resume (path: a NUL-terminated string, trim(name)//c_null_char) ----
     integer(c_int) function fvwb_checkpoint_save(h, path) bind(C, name='fvwb_checkpoint_save')
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: h
       character(kind=c_char), intent(in) :: path(*)
     end function fvwb_checkpoint_save
     integer(c_int) function fvwb_checkpoint_load(h, path) bind(C, name='fvwb_checkpoint_load')
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: h
       character(kind=c_char), intent(in) :: path(*)
     end function fvwb_checkpoint_load
     integer(c_int) function fvwb1d_checkpoint_save(h, path) bind(C, name='fvwb1d_checkpoint_save')
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: h
       character(kind=c_char), intent(in) :: path(*)
     end function fvwb1d_checkpoint_save
     integer(c_int) function fvwb1d_checkpoint_load(h, path) bind(C, name='fvwb1d_checkpoint_load')
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: h
       character(kind=c_char), intent(in) :: path(*)
     end function fvwb1d_checkpoint_load
     integer(c_int) function fvwb_dg_checkpoint_save(h, path) bind(C, name='fvwb_dg_checkpoint_save')
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: h
       character(kind=c_char), intent(in) :: path(*)
     end function fvwb_dg_checkpoint_save
     integer(c_int) function fvwb_dg_checkpoint_load(h, path) bind(C, name='fvwb_dg_checkpoint_load')
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: h
       character(kind=c_char), intent(in) :: path(*)
     end function fvwb_dg_checkpoint_load
     integer(c_int) function fvwb_dg1d_checkpoint_save(h, path) bind(C, name='fvwb_dg1d_checkpoint_save')
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: h
       character(kind=c_char), intent(in) :: path(*)
     end function fvwb_dg1d_checkpoint_save
     integer(c_int) function fvwb_dg1d_checkpoint_load(h, path) bind(C, name='fvwb_dg1d_checkpoint_load')
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: h
       character(kind=c_char), intent(in) :: path(*)
     end function fvwb_dg1d_checkpoint_load

     ! ---- 1D modal DG ----
     integer(c_int) function fvwb_dg1d_params_reference(p, nx) bind(C, name='fvwb_dg1d_params_reference')
       import :: c_int, c_int32_t, fvwb_dg1d_params
       type(fvwb_dg1d_params), intent(out) :: p
       integer(c_int32_t), value :: nx
     end function fvwb_dg1d_params_reference

     integer(c_int) function fvwb_dg1d_create(h, p) bind(C, name='fvwb_dg1d_create')
       import :: c_int, c_ptr, fvwb_dg1d_params
       type(c_ptr), intent(out) :: h
       type(fvwb_dg1d_params), intent(in) :: p
     end function fvwb_dg1d_create

     integer(c_int) function fvwb_dg1d_destroy(h) bind(C, name='fvwb_dg1d_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function fvwb_dg1d_destroy

     type(c_ptr) function fvwb_dg1d_last_error(h) bind(C, name='fvwb_dg1d_last_error')
       import :: c_ptr
       type(c_ptr), value :: h
     end function fvwb_dg1d_last_error

     ! drop-in for dg_with_source.f90:1749-2031 (compute_update_exact_delta(delta_u,u_eq,dudt))
     integer(c_int) function fvwb_dg1d_compute_update_exact_delta(h, delta_u, u_eq, dudt) &
          bind(C, name='fvwb_dg1d_compute_update_exact_delta')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: delta_u(*), u_eq(*)
       real(c_double), intent(out) :: dudt(*)
     end function fvwb_dg1d_compute_update_exact_delta

     ! the main program's 'RKi' loop (dg_with_source.f90:191-334), resident on the device
     integer(c_int) function fvwb_dg1d_upload(h, delta_u, u_eq, u_init) bind(C, name='fvwb_dg1d_upload')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: delta_u(*), u_eq(*), u_init(*)
     end function fvwb_dg1d_upload

     integer(c_int) function fvwb_dg1d_advance(h, nsteps, t_end) bind(C, name='fvwb_dg1d_advance')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: nsteps
       real(c_double), value :: t_end
     end function fvwb_dg1d_advance

     integer(c_int) function fvwb_dg1d_download(h, delta_u, u_init) bind(C, name='fvwb_dg1d_download')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: delta_u(*), u_init(*)
     end function fvwb_dg1d_download

     integer(c_int) function fvwb_dg1d_get_time(h, t, dt_last, steps) bind(C, name='fvwb_dg1d_get_time')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: t, dt_last
       integer(c_int32_t), intent(out) :: steps
     end function fvwb_dg1d_get_time

     ! ---- 2D residual distribution ----
     integer(c_int) function fvwb_rd_create(h, p) bind(C, name='fvwb_rd_create')
       import :: c_int, c_ptr, fvwb_rd_params
       type(c_ptr), intent(out) :: h
       type(fvwb_rd_params), intent(in) :: p
     end function fvwb_rd_create

     integer(c_int) function fvwb_rd_destroy(h) bind(C, name='fvwb_rd_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function fvwb_rd_destroy

     type(c_ptr) function fvwb_rd_last_error(h) bind(C, name='fvwb_rd_last_error')
       import :: c_ptr
       type(c_ptr), value :: h
     end function fvwb_rd_last_error

     ! drop-in for benchmark_rd.f90:612-745 (compute_update_rd(u,diff,dt,x,y,dudt))
     integer(c_int) function fvwb_rd_compute_update(h, u, diff, dt, dudt) bind(C, name='fvwb_rd_compute_update')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: u(*), diff(*)
       real(c_double), value :: dt
       real(c_double), intent(out) :: dudt(*)
     end function fvwb_rd_compute_update

     ! drop-in for benchmark_rd.f90:448-475 (compute_max_speed(u,speed_max))
     integer(c_int) function fvwb_rd_compute_max_speed(h, u, cmax) bind(C, name='fvwb_rd_compute_max_speed')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: u(*)
       real(c_double), intent(out) :: cmax
     end function fvwb_rd_compute_max_speed

     ! evolve with solver 'RD1' (benchmark_rd.f90:333-388), resident on the device
     integer(c_int) function fvwb_rd_upload(h, u) bind(C, name='fvwb_rd_upload')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: u(*)
     end function fvwb_rd_upload

     integer(c_int) function fvwb_rd_advance(h, nsteps, t_end) bind(C, name='fvwb_rd_advance')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: nsteps
       real(c_double), value :: t_end
     end function fvwb_rd_advance

     integer(c_int) function fvwb_rd_download(h, u) bind(C, name='fvwb_rd_download')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: u(*)
     end function fvwb_rd_download

     ! binary checkpoint of the resident state and clock (path NUL-terminated)
     integer(c_int) function fvwb_rd_checkpoint_save(h, path) bind(C, name='fvwb_rd_checkpoint_save')
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: h
       character(kind=c_char), intent(in) :: path(*)
     end function fvwb_rd_checkpoint_save
     integer(c_int) function fvwb_rd_checkpoint_load(h, path) bind(C, name='fvwb_rd_checkpoint_load')
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: h
       character(kind=c_char), intent(in) :: path(*)
     end function fvwb_rd_checkpoint_load

     integer(c_int) function fvwb_rd_get_time(h, t, dt_last, steps) bind(C, name='fvwb_rd_get_time')
       import :: c_int, c_int64_t, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: t, dt_last
       integer(c_int64_t), intent(out) :: steps
     end function fvwb_rd_get_time
     ! ---- 2D modal DG, configured P2 deviation scheme (include/fvwb.h) ----
     integer(c_int) function fvwb_dg_create(h, p) bind(C, name='fvwb_dg_create')
       import :: c_int, c_ptr, fvwb_dg_params
       type(c_ptr), intent(out) :: h
       type(fvwb_dg_params), intent(in) :: p
     end function fvwb_dg_create
     integer(c_int) function fvwb_dg_destroy(h) bind(C, name='fvwb_dg_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function fvwb_dg_destroy
     type(c_ptr) function fvwb_dg_last_error(h) bind(C, name='fvwb_dg_last_error')
       import :: c_ptr
       type(c_ptr), value :: h
     end function fvwb_dg_last_error
     ! the deviation form of compute_update (2d/benchmark_2d_dg.f90:1137-1479), modes (4,nx,ny,6)
     integer(c_int) function fvwb_dg_compute_residual(h, modes, dudt) bind(C, name='fvwb_dg_compute_residual')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: modes(*)
       real(c_double), intent(out) :: dudt(*)
     end function fvwb_dg_compute_residual
     integer(c_int) function fvwb_dg_upload(h, modes) bind(C, name='fvwb_dg_upload')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: modes(*)
     end function fvwb_dg_upload
     integer(c_int) function fvwb_dg_init_fourier(h, nterms, a, mk, nk, theta, amp) bind(C, name='fvwb_dg_init_fourier')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: nterms
       real(c_double), intent(in) :: a(*), mk(*), nk(*), theta(*)
       real(c_double), value :: amp
     end function fvwb_dg_init_fourier
     integer(c_int) function fvwb_dg_advance(h, nsteps, t_end) bind(C, name='fvwb_dg_advance')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: nsteps
       real(c_double), value :: t_end
     end function fvwb_dg_advance
     integer(c_int) function fvwb_dg_download(h, modes) bind(C, name='fvwb_dg_download')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: modes(*)
     end function fvwb_dg_download
     integer(c_int) function fvwb_dg_get_time(h, t, dt_last, steps_done) bind(C, name='fvwb_dg_get_time')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: t, dt_last
       integer(c_int32_t), intent(out) :: steps_done
     end function fvwb_dg_get_time

     ! ---- the reference's own 2D DG program (2d/benchmark_2d_dg.f90) ----
     integer(c_int) function fvwb_dgq_params_reference(p) bind(C, name='fvwb_dgq_params_reference')
       import :: c_int, fvwb_dgq_params
       type(fvwb_dgq_params), intent(out) :: p
     end function fvwb_dgq_params_reference
     ! 2d/legendre.f90:77-108 gl_quadrature(x_quad, w_quad, n)
     integer(c_int) function fvwb_dgq_gl_quadrature(n, x_quad, w_quad) bind(C, name='fvwb_dgq_gl_quadrature')
       import :: c_int, c_int32_t, c_double
       integer(c_int32_t), value :: n
       real(c_double), intent(out) :: x_quad(*), w_quad(*)
     end function fvwb_dgq_gl_quadrature
     integer(c_int) function fvwb_dgq_create(h, p) bind(C, name='fvwb_dgq_create')
       import :: c_int, c_ptr, fvwb_dgq_params
       type(c_ptr), intent(out) :: h
       type(fvwb_dgq_params), intent(in) :: p
     end function fvwb_dgq_create
     integer(c_int) function fvwb_dgq_destroy(h) bind(C, name='fvwb_dgq_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function fvwb_dgq_destroy
     type(c_ptr) function fvwb_dgq_last_error(h) bind(C, name='fvwb_dgq_last_error')
       import :: c_ptr
       type(c_ptr), value :: h
     end function fvwb_dgq_last_error
     ! :1137-1479 compute_update(delta_u, x, y, u_eq, dudt)
     integer(c_int) function fvwb_dgq_compute_update(h, delta_u, x, y, u_eq, dudt) &
          bind(C, name='fvwb_dgq_compute_update')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: delta_u(*), x(*), y(*), u_eq(*)
       real(c_double), intent(out) :: dudt(*)
     end function fvwb_dgq_compute_update
     ! :826-870 compute_max_speed(u, cs_max, v_xmax, v_ymax, speed_max)
     integer(c_int) function fvwb_dgq_compute_max_speed(h, u, cs_max, v_xmax, v_ymax, speed_max) &
          bind(C, name='fvwb_dgq_compute_max_speed')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: u(*)
       real(c_double), intent(out) :: cs_max, v_xmax, v_ymax, speed_max
     end function fvwb_dgq_compute_max_speed
     ! :1516-1555 apply_limiter(u), in place
     integer(c_int) function fvwb_dgq_apply_limiter(h, u) bind(C, name='fvwb_dgq_apply_limiter')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(inout) :: u(*)
     end function fvwb_dgq_apply_limiter
     ! :497-592 get_modes_from_nodes / get_nodes_from_modes
     integer(c_int) function fvwb_dgq_modes_from_nodes(h, nodes, modes) bind(C, name='fvwb_dgq_modes_from_nodes')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: nodes(*)
       real(c_double), intent(out) :: modes(*)
     end function fvwb_dgq_modes_from_nodes
     integer(c_int) function fvwb_dgq_nodes_from_modes(h, modes, nodes) bind(C, name='fvwb_dgq_nodes_from_modes')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: modes(*)
       real(c_double), intent(out) :: nodes(*)
     end function fvwb_dgq_nodes_from_modes
     ! :624-775 evolve, resident on the device
     integer(c_int) function fvwb_dgq_upload(h, delta_u, x, y) bind(C, name='fvwb_dgq_upload')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: delta_u(*), x(*), y(*)
     end function fvwb_dgq_upload
     integer(c_int) function fvwb_dgq_advance(h, nsteps, t_end) bind(C, name='fvwb_dgq_advance')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: nsteps
       real(c_double), value :: t_end
     end function fvwb_dgq_advance
     integer(c_int) function fvwb_dgq_download(h, delta_u) bind(C, name='fvwb_dgq_download')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: delta_u(*)
     end function fvwb_dgq_download
     integer(c_int) function fvwb_dgq_get_time(h, t, dt_last, steps_done) bind(C, name='fvwb_dgq_get_time')
       import :: c_int, c_int64_t, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: t, dt_last
       integer(c_int64_t), intent(out) :: steps_done
     end function fvwb_dgq_get_time
     ! :23-89 compute_error of host nodal arrays
     integer(c_int) function fvwb_dgq_compute_error(h, nodes, u_init, linf, l1, l2) &
          bind(C, name='fvwb_dgq_compute_error')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: nodes(*), u_init(*)
       real(c_double), intent(out) :: linf(4), l1(4), l2(4)
     end function fvwb_dgq_compute_error
  end interface

contains

  ! Copy a NUL-terminated C string into a Fortran string.
  function fvwb_message(cstr) result(msg)
    type(c_ptr), intent(in) :: cstr
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    if (.not. c_associated(cstr)) then
       msg = '(no message)'
       return
    end if
    n = int(c_strlen(cstr))
    call c_f_pointer(cstr, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
       msg(i:i) = chars(i)
    end do
  end function fvwb_message

  ! Stop with the library's message when a call on a single-problem handle failed.
  subroutine fvwb_check(status, h, what)
    integer(c_int), intent(in) :: status
    type(c_ptr), intent(in) :: h
    character(len=*), intent(in) :: what
    if (status /= FVWB_OK) then
       write(0, '(a,a,a,i0,a,a)') 'fvwb: ', what, ' failed with status ', status, ': ', &
            fvwb_message(fvwb_last_error(h))
       error stop 1
    end if
  end subroutine fvwb_check

  ! Same for a batch handle.
  subroutine fvwb_batch_check(status, h, what)
    integer(c_int), intent(in) :: status
    type(c_ptr), intent(in) :: h
    character(len=*), intent(in) :: what
    if (status /= FVWB_OK) then
       write(0, '(a,a,a,i0,a,a)') 'fvwb: ', what, ' failed with status ', status, ': ', &
            fvwb_message(fvwb_batch_last_error(h))
       error stop 1
    end if
  end subroutine fvwb_batch_check

  ! Same for a 1D handle.
  subroutine fvwb1d_check(status, h, what)
    integer(c_int), intent(in) :: status
    type(c_ptr), intent(in) :: h
    character(len=*), intent(in) :: what
    if (status /= FVWB_OK) then
       write(0, '(a,a,a,i0,a,a)') 'fvwb: ', what, ' failed with status ', status, ': ', &
            fvwb_message(fvwb1d_last_error(h))
       error stop 1
    end if
  end subroutine fvwb1d_check

  ! Same for a 1D DG handle.
  subroutine fvwb_dg1d_check(status, h, what)
    integer(c_int), intent(in) :: status
    type(c_ptr), intent(in) :: h
    character(len=*), intent(in) :: what
    if (status /= FVWB_OK) then
       write(0, '(a,a,a,i0,a,a)') 'fvwb: ', what, ' failed with status ', status, ': ', &
            fvwb_message(fvwb_dg1d_last_error(h))
       error stop 1
    end if
  end subroutine fvwb_dg1d_check

  subroutine fvwb_rd_check(status, h, what)
    integer(c_int), intent(in) :: status
    type(c_ptr), intent(in) :: h
    character(len=*), intent(in) :: what
    if (status /= FVWB_OK) then
       write(0, '(a,a,a,i0,a,a)') 'fvwb: ', what, ' failed with status ', status, ': ', &
            fvwb_message(fvwb_rd_last_error(h))
       error stop 1
    end if
  end subroutine fvwb_rd_check

  subroutine fvwb_dg_check(status, h, what)
    integer(c_int), intent(in) :: status
    type(c_ptr), intent(in) :: h
    character(len=*), intent(in) :: what
    if (status /= FVWB_OK) then
       write(0, '(a,a,a,i0,a,a)') 'fvwb: ', what, ' failed with status ', status, ': ', &
            fvwb_message(fvwb_dg_last_error(h))
       error stop 1
    end if
  end subroutine fvwb_dg_check

  subroutine fvwb_dgq_check(status, h, what)
    integer(c_int), intent(in) :: status
    type(c_ptr), intent(in) :: h
    character(len=*), intent(in) :: what
    if (status /= FVWB_OK) then
       write(0, '(a,a,a,i0,a,a)') 'fvwb: ', what, ' failed with status ', status, ': ', &
            fvwb_message(fvwb_dgq_last_error(h))
       error stop 1
    end if
  end subroutine fvwb_dgq_check

end module fvwb_iface
