! Drop-in replacement for the residual of the reference's 1D DG program.
!
! compute_update_exact_delta(delta_u, u_eq, dudt) keeps the external signature
! of dg_with_source.f90:1749-1753 (implicit interface, explicit-shape
! real(kind=8) arrays (1:nvar,1:n,1:nx), sizes and switches from module
! dg_commons).  The first call creates one library handle from dg_commons (nx,
! n, riemann, source, gamma, boxlen); fvwb_dg1d_dropin_close() releases it.  A
! failing call stops the program with the library's message.
module fvwb_dropin_dg1d
  use iso_c_binding
  use fvwb_iface
  implicit none
  private
  type(c_ptr), save :: handle = c_null_ptr
  public :: fvwb_dg1d_dropin_handle, fvwb_dg1d_dropin_close

contains

  function fvwb_dg1d_dropin_handle() result(h)
    use dg_commons, only: nx, n, riemann, source, gamma, boxlen
    type(c_ptr) :: h
    type(fvwb_dg1d_params) :: p
    if (.not. c_associated(handle)) then
       call fvwb_dg1d_check(fvwb_dg1d_params_reference(p, int(nx, c_int32_t)), c_null_ptr, &
            'fvwb_dg1d_params_reference')
       p%n = int(n, c_int32_t)
       p%riemann = int(riemann, c_int32_t)
       p%source = int(source, c_int32_t)
       p%gamma = gamma
       p%boxlen = boxlen
       call fvwb_dg1d_check(fvwb_dg1d_create(handle, p), c_null_ptr, 'fvwb_dg1d_create')
    end if
    h = handle
  end function fvwb_dg1d_dropin_handle

  subroutine fvwb_dg1d_dropin_close()
    if (c_associated(handle)) then
       call fvwb_dg1d_check(fvwb_dg1d_destroy(handle), c_null_ptr, 'fvwb_dg1d_destroy')
       handle = c_null_ptr
    end if
  end subroutine fvwb_dg1d_dropin_close

end module fvwb_dropin_dg1d

! dudt of the well-balanced deviation-form DG residual (dg_with_source.f90:1749-2031)
subroutine compute_update_exact_delta(delta_u, u_eq, dudt)
  use dg_commons, only: nvar, n, nx
  use fvwb_iface, only: fvwb_dg1d_compute_update_exact_delta, fvwb_dg1d_check
  use fvwb_dropin_dg1d, only: fvwb_dg1d_dropin_handle
  use iso_c_binding, only: c_ptr
  implicit none
  real(kind=8), dimension(1:nvar, 1:n, 1:nx) :: delta_u, u_eq, dudt
  type(c_ptr) :: h
  h = fvwb_dg1d_dropin_handle()
  call fvwb_dg1d_check(fvwb_dg1d_compute_update_exact_delta(h, delta_u, u_eq, dudt), h, &
       'compute_update_exact_delta')
end subroutine compute_update_exact_delta
