! Drop-in replacements for the reference's 2D per-stage routines.
!
! compute_update_exact(u, w_eq, dudt) and compute_max_speed(u, cmax) keep the
! reference's external-procedure signatures (benchmark_2d.f90:465-468 and
! :264-268: implicit interface, explicit-shape real(kind=8) arrays
! (1:nvar,1:nx,1:ny), sizes and constants taken from module parameters_2d), so a
! driver written against them links against these instead and every stage runs
! in the HIP kernels of libfvwb.  The first call creates one library handle from
! parameters_2d (nx, ny, nequilibrium, gamma, box, cfl); fvwb_dropin_close()
! releases it.  A failing call stops the program with the library's message.
module fvwb_dropin_2d
  use iso_c_binding
  use fvwb_iface
  implicit none
  private
  type(c_ptr), save :: handle = c_null_ptr
  public :: fvwb_dropin_handle, fvwb_dropin_close

contains

  function fvwb_dropin_handle() result(h)
    use parameters_2d, only: nx, ny, nequilibrium, gamma, boxlen_x, boxlen_y, cfl
    type(c_ptr) :: h
    type(fvwb_params) :: p
    if (.not. c_associated(handle)) then
       call fvwb_check(fvwb_params_reference_2d(p, int(nx, c_int32_t), int(ny, c_int32_t), &
            int(nequilibrium, c_int32_t)), c_null_ptr, 'fvwb_params_reference_2d')
       p%gamma = gamma
       p%boxlen_x = boxlen_x
       p%boxlen_y = boxlen_y
       p%cfl = cfl
       call fvwb_check(fvwb_create(handle, p), c_null_ptr, 'fvwb_create')
    end if
    h = handle
  end function fvwb_dropin_handle

  subroutine fvwb_dropin_close()
    if (c_associated(handle)) then
       call fvwb_check(fvwb_destroy(handle), c_null_ptr, 'fvwb_destroy')
       handle = c_null_ptr
    end if
  end subroutine fvwb_dropin_close

end module fvwb_dropin_2d

! dudt = exact well-balanced deviation residual of u about the equilibrium whose
! primitive centre values are w_eq (benchmark_2d.f90:465-618); ring cells zero.
subroutine compute_update_exact(u, w_eq, dudt)
  use parameters_2d, only: nvar, nx, ny
  use fvwb_iface, only: fvwb_compute_update_exact, fvwb_check
  use fvwb_dropin_2d, only: fvwb_dropin_handle
  use iso_c_binding, only: c_ptr
  implicit none
  real(kind=8), dimension(1:nvar, 1:nx, 1:ny) :: u, w_eq, dudt
  type(c_ptr) :: h
  h = fvwb_dropin_handle()
  call fvwb_check(fvwb_compute_update_exact(h, u, w_eq, dudt), h, 'compute_update_exact')
end subroutine compute_update_exact

! The plain, non-well-balanced residual (benchmark_2d.f90:370-463).
subroutine compute_update(u, w_eq, dudt)
  use parameters_2d, only: nvar, nx, ny
  use fvwb_iface, only: fvwb_compute_update, fvwb_check
  use fvwb_dropin_2d, only: fvwb_dropin_handle
  use iso_c_binding, only: c_ptr
  implicit none
  real(kind=8), dimension(1:nvar, 1:nx, 1:ny) :: u, w_eq, dudt
  type(c_ptr) :: h
  h = fvwb_dropin_handle()
  call fvwb_check(fvwb_compute_update(h, u, w_eq, dudt), h, 'compute_update')
end subroutine compute_update

! cmax = max over cells of |v| + c (benchmark_2d.f90:264-279).
subroutine compute_max_speed(u, cmax)
  use parameters_2d, only: nvar, nx, ny
  use fvwb_iface, only: fvwb_compute_max_speed, fvwb_check
  use fvwb_dropin_2d, only: fvwb_dropin_handle
  use iso_c_binding, only: c_ptr
  implicit none
  real(kind=8), dimension(1:nvar, 1:nx, 1:ny) :: u
  real(kind=8) :: cmax
  type(c_ptr) :: h
  h = fvwb_dropin_handle()
  call fvwb_check(fvwb_compute_max_speed(h, u, cmax), h, 'compute_max_speed')
end subroutine compute_max_speed
