! Drop-ins for the reference's 2D DG program (2d/benchmark_2d_dg.f90) on the
! GPU: external procedures with the reference's names and dummy arguments,
! reading sizes and switches from module parameters_dg_2d, each forwarding to
! libfvwb's fvwb_dgq_* entry points (FP64, the reference's results bit for
! bit).  Link them in place of the reference's own procedures.  One handle is
! created on first use and kept until fvwb_dgq_dropin_close.
module fvwb_dropin_dg2d
  use iso_c_binding
  use fvwb_iface
  use parameters_dg_2d
  implicit none
  private
  type(c_ptr), save :: handle = c_null_ptr
  public :: fvwb_dgq_dropin_params, fvwb_dgq_dropin_handle, fvwb_dgq_dropin_close

contains

  ! the module's switches as the library's parameter block
  function fvwb_dgq_dropin_params() result(p)
    type(fvwb_dgq_params) :: p
    p%nx = nx
    p%ny = ny
    if (mx /= my) error stop 'fvwb: the drop-ins need mx == my'
    p%order = mx
    select case (flux_type)
    case ('llf1')
       p%flux_type = FVWB_DGQ_FLUX_LLF1
    case ('hll2')
       p%flux_type = FVWB_DGQ_FLUX_HLL2
    case ('hllc')
       p%flux_type = FVWB_DGQ_FLUX_HLLC
    case default                 ! no branch of compute_num_flux: zero face flux
       p%flux_type = FVWB_DGQ_FLUX_NONE
    end select
    p%source = source
    p%grad_phi_case = grad_phi_case
    p%bc = bc
    p%ninit = ninit
    select case (solver)
    case ('RK4')
       p%solver = FVWB_DGQ_SOLVER_RK4
    case ('SS4')
       p%solver = FVWB_DGQ_SOLVER_SS4
    case ('EQL')
       p%solver = FVWB_DGQ_SOLVER_EQL
    case ('DEB')
       p%solver = FVWB_DGQ_SOLVER_DEB
    case default
       error stop 'fvwb: solver must be RK4, SS4, EQL or DEB'
    end select
    if (use_limiter .and. limiter_type /= 'ONP') error stop 'fvwb: the drop-ins implement limiter_type ONP'
    p%use_limiter = merge(1_c_int32_t, 0_c_int32_t, use_limiter)
    p%gamma = gamma
    p%boxlen_x = boxlen_x
    p%boxlen_y = boxlen_y
    p%cfl = cfl
    p%eps = eps
  end function fvwb_dgq_dropin_params

  function fvwb_dgq_dropin_handle() result(h)
    type(c_ptr) :: h
    if (.not. c_associated(handle)) &
         call fvwb_dgq_check(fvwb_dgq_create(handle, fvwb_dgq_dropin_params()), c_null_ptr, 'fvwb_dgq_create')
    h = handle
  end function fvwb_dgq_dropin_handle

  subroutine fvwb_dgq_dropin_close()
    if (c_associated(handle)) call fvwb_dgq_check(fvwb_dgq_destroy(handle), c_null_ptr, 'fvwb_dgq_destroy')
    handle = c_null_ptr
  end subroutine fvwb_dgq_dropin_close

end module fvwb_dropin_dg2d

! 2d/benchmark_2d_dg.f90:1137-1479
subroutine compute_update(delta_u, x, y, u_eq, dudt)
  use iso_c_binding
  use parameters_dg_2d
  use fvwb_iface
  use fvwb_dropin_dg2d
  implicit none
  real(kind=8), dimension(1:nvar, 1:nx, 1:ny, 1:mx, 1:my) :: delta_u, u_eq, dudt
  real(kind=8), dimension(1:nx, 1:ny, 1:mx, 1:my) :: x, y
  type(c_ptr) :: h
  h = fvwb_dgq_dropin_handle()
  call fvwb_dgq_check(fvwb_dgq_compute_update(h, delta_u, x, y, u_eq, dudt), h, 'fvwb_dgq_compute_update')
end subroutine compute_update

! :826-870
subroutine compute_max_speed(u, cs_max, v_xmax, v_ymax, speed_max)
  use iso_c_binding
  use parameters_dg_2d
  use fvwb_iface
  use fvwb_dropin_dg2d
  implicit none
  real(kind=8), dimension(1:nvar, 1:nx, 1:ny) :: u
  real(kind=8) :: cs_max, v_xmax, v_ymax, speed_max
  type(c_ptr) :: h
  h = fvwb_dgq_dropin_handle()
  call fvwb_dgq_check(fvwb_dgq_compute_max_speed(h, u, cs_max, v_xmax, v_ymax, speed_max), h, &
       'fvwb_dgq_compute_max_speed')
end subroutine compute_max_speed

! :1516-1555 (limiter_type 'ONP': compute_positivity of 2d/limiters.f90:478-654)
subroutine apply_limiter(u)
  use iso_c_binding
  use parameters_dg_2d
  use fvwb_iface
  use fvwb_dropin_dg2d
  implicit none
  real(kind=8), dimension(1:nvar, 1:nx, 1:ny, 1:mx, 1:my) :: u
  type(c_ptr) :: h
  h = fvwb_dgq_dropin_handle()
  call fvwb_dgq_check(fvwb_dgq_apply_limiter(h, u), h, 'fvwb_dgq_apply_limiter')
end subroutine apply_limiter

! :497-542
subroutine get_modes_from_nodes(nodes, u, size_x, size_y, order_x, order_y)
  use iso_c_binding
  use parameters_dg_2d
  use fvwb_iface
  use fvwb_dropin_dg2d
  implicit none
  integer :: size_x, size_y, order_x, order_y
  real(kind=8), dimension(1:nvar, 1:size_x, 1:size_y, 1:order_x, 1:order_y) :: nodes, u
  type(c_ptr) :: h
  if (size_x /= nx .or. size_y /= ny .or. order_x /= mx .or. order_y /= my) &
       error stop 'fvwb: get_modes_from_nodes on another grid than parameters_dg_2d'
  h = fvwb_dgq_dropin_handle()
  call fvwb_dgq_check(fvwb_dgq_modes_from_nodes(h, nodes, u), h, 'fvwb_dgq_modes_from_nodes')
end subroutine get_modes_from_nodes

! :544-592
subroutine get_nodes_from_modes(u, nodes, size_x, size_y, order_x, order_y)
  use iso_c_binding
  use parameters_dg_2d
  use fvwb_iface
  use fvwb_dropin_dg2d
  implicit none
  integer :: size_x, size_y, order_x, order_y
  real(kind=8), dimension(1:nvar, 1:size_x, 1:size_y, 1:order_x, 1:order_y) :: u, nodes
  type(c_ptr) :: h
  if (size_x /= nx .or. size_y /= ny .or. order_x /= mx .or. order_y /= my) &
       error stop 'fvwb: get_nodes_from_modes on another grid than parameters_dg_2d'
  h = fvwb_dgq_dropin_handle()
  call fvwb_dgq_check(fvwb_dgq_nodes_from_modes(h, u, nodes), h, 'fvwb_dgq_nodes_from_modes')
end subroutine get_nodes_from_modes

! 2d/legendre.f90:77-108
subroutine gl_quadrature(x_quad, w_quad, n)
  use iso_c_binding
  use fvwb_iface
  implicit none
  integer :: n
  real(kind=8), dimension(1:n) :: x_quad, w_quad
  call fvwb_dgq_check(fvwb_dgq_gl_quadrature(int(n, c_int32_t), x_quad, w_quad), c_null_ptr, 'fvwb_dgq_gl_quadrature')
end subroutine gl_quadrature

! :23-89: the norms against the program's initial conditions, printed as the
! reference prints them (the device reduction's sums round differently from
! the reference's serial loop, so the last digits may differ)
subroutine compute_error(u, x, y, t, u_anal)
  use iso_c_binding
  use parameters_dg_2d
  use fvwb_iface
  use fvwb_dropin_dg2d
  implicit none
  real(kind=8), dimension(1:nx, 1:ny, 1:mx, 1:my) :: x, y
  real(kind=8), dimension(1:nvar, 1:nx, 1:ny, 1:mx, 1:my) :: u, u_init, u_anal
  real(kind=8) :: t
  real(c_double) :: linf(4), l1(4), l2(4)
  type(c_ptr) :: h
  external :: get_initial_conditions
  call get_initial_conditions(x, y, u_init, nx, ny, mx, my)
  h = fvwb_dgq_dropin_handle()
  call fvwb_dgq_check(fvwb_dgq_compute_error(h, u, u_init, linf, l1, l2), h, 'fvwb_dgq_compute_error')
  print *, 'Grid size, Order', nx, mx
  print *, 'maxerror rho', linf(1)
  print *, 'maxerror velx', linf(2)
  print *, 'maxerror vely', linf(3)
  print *, 'maxerror energy', linf(4)
  print *, 'l1error rho', l1(1)
  print *, 'l1error velx', l1(2)
  print *, 'l1error vely', l1(3)
  print *, 'l1error energy', l1(4)
  print *, 'l2error rho', l2(1)
  print *, 'l2error velx', l2(2)
  print *, 'l2error vely', l2(3)
  print *, 'l2error energy', l2(4)
end subroutine compute_error
