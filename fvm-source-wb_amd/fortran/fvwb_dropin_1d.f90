! Drop-in replacements for the reference's 1D per-stage routines.
!
! compute_update(u, w_eq, dudt), compute_update_sr(u, w_eq, dudt),
! compute_update_fvm(u, w_eq, dudt) and compute_max_speed(u, cmax) keep the external
! signatures of benchmark_1d.f90:263-266, :553-556, :454-457 and :157-162 (implicit interface,
! explicit-shape real(kind=8) arrays (1:nvar,1:nx), sizes and constants from
! module parameters).  The first call creates one library handle from
! parameters (nx, nequilibrium, gamma, boxlen, bc); fvwb1d_dropin_close()
! releases it.  A failing call stops the program with the library's message.
module fvwb_dropin_1d
  use iso_c_binding
  use fvwb_iface
  implicit none
  private
  type(c_ptr), save :: handle = c_null_ptr
  public :: fvwb1d_dropin_handle, fvwb1d_dropin_close

contains

  function fvwb1d_dropin_handle() result(h)
    use parameters, only: nx, nequilibrium, gamma, boxlen, bc
    type(c_ptr) :: h
    type(fvwb1d_params) :: p(1)
    if (.not. c_associated(handle)) then
       call fvwb1d_check(fvwb1d_params_reference(p(1), int(nx, c_int32_t), int(nequilibrium, c_int32_t)), &
            c_null_ptr, 'fvwb1d_params_reference')
       p(1)%gamma = gamma
       p(1)%boxlen = boxlen
       p(1)%bc = int(bc, c_int32_t)
       call fvwb1d_check(fvwb1d_create(handle, 1_c_int32_t, p), c_null_ptr, 'fvwb1d_create')
    end if
    h = handle
  end function fvwb1d_dropin_handle

  subroutine fvwb1d_dropin_close()
    if (c_associated(handle)) then
       call fvwb1d_check(fvwb1d_destroy(handle), c_null_ptr, 'fvwb1d_destroy')
       handle = c_null_ptr
    end if
  end subroutine fvwb1d_dropin_close

end module fvwb_dropin_1d

! dudt of the exact deviation residual (benchmark_1d.f90:263-377)
subroutine compute_update(u, w_eq, dudt)
  use parameters, only: nvar, nx
  use fvwb_iface, only: fvwb1d_compute_update, fvwb1d_check
  use fvwb_dropin_1d, only: fvwb1d_dropin_handle
  use iso_c_binding, only: c_ptr
  implicit none
  real(kind=8), dimension(1:nvar, 1:nx) :: u, w_eq, dudt
  type(c_ptr) :: h
  h = fvwb1d_dropin_handle()
  call fvwb1d_check(fvwb1d_compute_update(h, u, w_eq, dudt), h, 'compute_update')
end subroutine compute_update

! dudt of the hydrostatic-reconstruction residual, solver 'WB1'
! (benchmark_1d.f90:553-747); w_eq is unused, as in the reference
subroutine compute_update_sr(u, w_eq, dudt)
  use parameters, only: nvar, nx
  use fvwb_iface, only: fvwb1d_compute_update_sr, fvwb1d_check
  use fvwb_dropin_1d, only: fvwb1d_dropin_handle
  use iso_c_binding, only: c_ptr
  implicit none
  real(kind=8), dimension(1:nvar, 1:nx) :: u, w_eq, dudt
  type(c_ptr) :: h
  h = fvwb1d_dropin_handle()
  call fvwb1d_check(fvwb1d_compute_update_sr(h, u, dudt), h, 'compute_update_sr')
end subroutine compute_update_sr

! dudt of the plain residual, solver 'FVM' (benchmark_1d.f90:454-549); w_eq is
! unused, as in the reference
subroutine compute_update_fvm(u, w_eq, dudt)
  use parameters, only: nvar, nx
  use fvwb_iface, only: fvwb1d_compute_update_fvm, fvwb1d_check
  use fvwb_dropin_1d, only: fvwb1d_dropin_handle
  use iso_c_binding, only: c_ptr
  implicit none
  real(kind=8), dimension(1:nvar, 1:nx) :: u, w_eq, dudt
  type(c_ptr) :: h
  h = fvwb1d_dropin_handle()
  call fvwb1d_check(fvwb1d_compute_update_fvm(h, u, dudt), h, 'compute_update_fvm')
end subroutine compute_update_fvm

! cmax = max over cells of |v| + c (benchmark_1d.f90:157-169)
subroutine compute_max_speed(u, cmax)
  use parameters, only: nvar, nx
  use fvwb_iface, only: fvwb1d_compute_max_speed, fvwb1d_check
  use fvwb_dropin_1d, only: fvwb1d_dropin_handle
  use iso_c_binding, only: c_ptr
  implicit none
  real(kind=8), dimension(1:nvar, 1:nx) :: u
  real(kind=8) :: cmax
  type(c_ptr) :: h
  h = fvwb1d_dropin_handle()
  call fvwb1d_check(fvwb1d_compute_max_speed(h, u, cmax), h, 'compute_max_speed')
end subroutine compute_max_speed
