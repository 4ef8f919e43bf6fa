! Drop-in replacements for the per-stage routines of the residual-distribution
! program benchmark_rd.f90.
!
! compute_update_rd(u, diff, dt, x, y, dudt) and compute_max_speed(u, speed_max)
! keep the external signatures of benchmark_rd.f90:612-620 and :448-451
! (implicit interface, explicit-shape real(kind=8) arrays
! (1:nvar,1:nx,1:ny,1:mx,1:my), sizes and switches from module parameters_rd).
! The first call creates one library handle from parameters_rd; fvwb_rd_dropin_close()
! releases it.  A failing call stops the program with the library's message.
module fvwb_dropin_rd
  use iso_c_binding
  use fvwb_iface
  implicit none
  private
  type(c_ptr), save :: handle = c_null_ptr
  public :: fvwb_rd_dropin_handle, fvwb_rd_dropin_close, fvwb_rd_dropin_params

contains

  function fvwb_rd_dropin_params() result(p)
    use parameters_rd, only: nx, ny, bc, gamma, boxlen_x, boxlen_y, cfl
    type(fvwb_rd_params) :: p
    p%nx = int(nx, c_int32_t)
    p%ny = int(ny, c_int32_t)
    p%bc = int(bc, c_int32_t)
    p%pad = 0
    p%gamma = gamma
    p%boxlen_x = boxlen_x
    p%boxlen_y = boxlen_y
    p%cfl = cfl
  end function fvwb_rd_dropin_params

  function fvwb_rd_dropin_handle() result(h)
    type(c_ptr) :: h
    if (.not. c_associated(handle)) &
         call fvwb_rd_check(fvwb_rd_create(handle, fvwb_rd_dropin_params()), c_null_ptr, 'fvwb_rd_create')
    h = handle
  end function fvwb_rd_dropin_handle

  subroutine fvwb_rd_dropin_close()
    if (c_associated(handle)) then
       call fvwb_rd_check(fvwb_rd_destroy(handle), c_null_ptr, 'fvwb_rd_destroy')
       handle = c_null_ptr
    end if
  end subroutine fvwb_rd_dropin_close

end module fvwb_dropin_rd

! dudt of the distributed residuals (benchmark_rd.f90:612-745); x and y only
! feed get_source's unused coordinates
subroutine compute_update_rd(u, diff, dt, x, y, dudt)
  use parameters_rd, only: nvar, nx, ny, mx, my
  use fvwb_iface, only: fvwb_rd_compute_update, fvwb_rd_check
  use fvwb_dropin_rd, only: fvwb_rd_dropin_handle
  use iso_c_binding, only: c_ptr
  implicit none
  real(kind=8), intent(in) :: dt
  real(kind=8), dimension(1:nvar, 1:nx, 1:ny, 1:mx, 1:my), intent(in) :: u, diff
  real(kind=8), dimension(1:nx, 1:ny, 1:mx, 1:my), intent(in) :: x, y
  real(kind=8), dimension(1:nvar, 1:nx, 1:ny, 1:mx, 1:my), intent(out) :: dudt
  type(c_ptr) :: h
  h = fvwb_rd_dropin_handle()
  call fvwb_rd_check(fvwb_rd_compute_update(h, u, diff, dt, dudt), h, 'compute_update_rd')
end subroutine compute_update_rd

! speed_max = max over nodes of |v| + c (benchmark_rd.f90:448-475)
subroutine compute_max_speed(u, speed_max)
  use parameters_rd, only: nvar, nx, ny, mx, my
  use fvwb_iface, only: fvwb_rd_compute_max_speed, fvwb_rd_check
  use fvwb_dropin_rd, only: fvwb_rd_dropin_handle
  use iso_c_binding, only: c_ptr
  implicit none
  real(kind=8), dimension(1:nvar, 1:nx, 1:ny, 1:mx, 1:my) :: u
  real(kind=8) :: speed_max
  type(c_ptr) :: h
  h = fvwb_rd_dropin_handle()
  call fvwb_rd_check(fvwb_rd_compute_max_speed(h, u, speed_max), h, 'compute_max_speed')
end subroutine compute_max_speed
