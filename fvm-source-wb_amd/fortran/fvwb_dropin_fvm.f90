! Drop-in replacements for the per-stage routines of the stand-alone 1D program
! fvm.f90.
!
! compute_update(u, dudt) and compute_max_speed(u, cmax) keep the external
! signatures of fvm.f90:188-191 and :320-323 (implicit interface, explicit-shape
! real(kind=8) arrays (1:nvar,1:nx), sizes and switches from module fvm_commons).
! The first call creates one library handle from fvm_commons (nx, gamma, boxlen,
! bc); fvwb_fvm_dropin_close() releases it.  A failing call stops the program
! with the library's message.
module fvwb_dropin_fvm
  use iso_c_binding
  use fvwb_iface
  implicit none
  private
  type(c_ptr), save :: handle = c_null_ptr
  public :: fvwb_fvm_dropin_handle, fvwb_fvm_dropin_close

contains

  function fvwb_fvm_dropin_handle() result(h)
    use fvm_commons, only: nx, gamma, boxlen, bc
    type(c_ptr) :: h
    type(fvwb1d_params) :: p(1)
    if (.not. c_associated(handle)) then
       call fvwb1d_check(fvwb1d_params_reference(p(1), int(nx, c_int32_t), 2_c_int32_t), &
            c_null_ptr, 'fvwb1d_params_reference')
       p(1)%gamma = gamma
       p(1)%boxlen = boxlen
       p(1)%bc = int(bc, c_int32_t)
       call fvwb1d_check(fvwb1d_create(handle, 1_c_int32_t, p), c_null_ptr, 'fvwb1d_create')
    end if
    h = handle
  end function fvwb_fvm_dropin_handle

  subroutine fvwb_fvm_dropin_close()
    if (c_associated(handle)) then
       call fvwb1d_check(fvwb1d_destroy(handle), c_null_ptr, 'fvwb1d_destroy')
       handle = c_null_ptr
    end if
  end subroutine fvwb_fvm_dropin_close

end module fvwb_dropin_fvm

! dudt of the plain Rusanov residual with the optional gravity source (fvm.f90:188-251)
subroutine compute_update(u, dudt)
  use fvm_commons, only: nvar, nx, source
  use fvwb_iface, only: fvwb1d_fvm_compute_update, fvwb1d_check
  use fvwb_dropin_fvm, only: fvwb_fvm_dropin_handle
  use iso_c_binding, only: c_ptr, c_int32_t
  implicit none
  real(kind=8), dimension(1:nvar, 1:nx) :: u, dudt
  type(c_ptr) :: h
  h = fvwb_fvm_dropin_handle()
  call fvwb1d_check(fvwb1d_fvm_compute_update(h, int(source, c_int32_t), u, dudt), h, 'compute_update')
end subroutine compute_update

! cmax = max over cells of |v| + c (fvm.f90:320-336)
subroutine compute_max_speed(u, cmax)
  use fvm_commons, only: nvar, nx
  use fvwb_iface, only: fvwb1d_compute_max_speed, fvwb1d_check
  use fvwb_dropin_fvm, only: fvwb_fvm_dropin_handle
  use iso_c_binding, only: c_ptr
  implicit none
  real(kind=8), dimension(1:nvar, 1:nx) :: u
  real(kind=8) :: cmax
  type(c_ptr) :: h
  h = fvwb_fvm_dropin_handle()
  call fvwb1d_check(fvwb1d_compute_max_speed(h, u, cmax), h, 'compute_max_speed')
end subroutine compute_max_speed
