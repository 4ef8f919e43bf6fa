! dg1d_fvwb — the reference's 1D DG program (dg_with_source.f90, integrator
! 'RKi') run on MI355X through libfvwb.
!
! The set-up of the main program (:27-31, 148-167): Gauss points, the
! equilibrium rho = p = exp(-x) at them, the initial state of condinit cases 7
! and 8 (:1116-1124) and its modes; then the 'RKi' loop (:191-199, 264-288,
! 303-334) with dt = 0.9*dx/cmax/(2n+1) from the first Gauss point of the
! reconstruction; then the final reconstruction written as recon99999.dat
! (x, p - p_eq at the first Gauss point) and hydro99999.dat (x, p), and the
! printed 'error' max |p - p_eq| (:337-406).
!
! usage: dg1d_fvwb [stage|resident] [tend] [prefix]
!   stage     the loop above on the host; every residual is the drop-in
!             compute_update_exact_delta; prints 'time=' iter t dt every step
!   resident  the whole loop on the device (fvwb_dg1d_advance)
!   prefix    the files become <prefix>recon99999.dat and <prefix>hydro99999.dat,
!             and <prefix>.bin holds int32 nx, n, nvar, steps; real(8) t, dt;
!             delta_u, u_init
program dg1d_fvwb
  use iso_c_binding
  use dg_commons
  use fvwb_iface
  use fvwb_dropin_dg1d
  use fvwb_basis1d
  implicit none
  real(kind=8), dimension(1:nvar, 1:n, 1:nx) :: delta_u, u_eq, uinit
  real(kind=8), dimension(1:nvar) :: ww, ww_e
  character(len=512) :: arg, mode, prefix
  integer :: nargs, ios, icell, i, j, iter, unit
  real(kind=8) :: dx, xcell, x_quad, t, dt, error

  mode = 'stage'
  prefix = ''
  nargs = command_argument_count()
  if (nargs >= 1) call get_command_argument(1, mode)
  if (nargs >= 2) then
     call get_command_argument(2, arg)
     read(arg, *, iostat=ios) tend
     if (ios /= 0) error stop 'dg1d_fvwb: tend must be a number'
  end if
  if (nargs >= 3) call get_command_argument(3, prefix)

  call gauss1d(chsi_quad, w_quad, nquad)
  dx = boxlen / dble(nx)
  delta_u(:, :, :) = 0.0
  do icell = 1, nx
     xcell = (dble(icell) - 0.5) * dx
     do j = 1, nquad
        x_quad = xcell + dx / 2.0 * chsi_quad(j)
        call initial_state(x_quad, uinit(1:nvar, j, icell))
        ww_e = [exp(-x_quad), 0d0, exp(-x_quad)]
        call conservative(ww_e, u_eq(1:nvar, j, icell))
     end do
     do i = 1, n
        do j = 1, nquad
           delta_u(1:nvar, i, icell) = delta_u(1:nvar, i, icell) + 0.5 * &
                & (uinit(1:nvar, j, icell) - u_eq(1:nvar, j, icell)) * basis1d(chsi_quad(j), i - 1) * w_quad(j)
        end do
     end do
  end do

  select case (trim(mode))
  case ('stage')
     call evolve_stages(t, dt, iter)
  case ('resident')
     call evolve_resident(t, dt, iter)
  case default
     error stop 'dg1d_fvwb: mode must be stage or resident'
  end select

  ! final state at the first Gauss point (:337-406)
  error = 0.0
  open(newunit=unit, file=trim(prefix)//'recon99999.dat', form='formatted', status='replace')
  do icell = 1, nx
     xcell = (dble(icell) - 0.5) * dx + dx / 2.0 * chsi_quad(1)
     ww_e = [exp(-xcell), 0d0, exp(-xcell)]
     call primitive(uinit(1:nvar, 1, icell), ww)
     write(unit, '(7(1PE12.5,1X))') xcell, ww(3) - ww_e(3)
     if (error < abs(ww(3) - ww_e(3))) error = abs(ww(3) - ww_e(3))
  end do
  close(unit)
  open(newunit=unit, file=trim(prefix)//'hydro99999.dat', form='formatted', status='replace')
  do icell = 1, nx
     xcell = (dble(icell) - 0.5) * dx + dx / 2.0 * chsi_quad(1)
     call primitive(uinit(1:nvar, 1, icell), ww)
     write(unit, '(7(1PE12.5,1X))') xcell, ww(3)
  end do
  close(unit)
  write(*, *) 'error', error
  write(*, '(a,a,a,i0,a,i0,a,i0,a,1pe24.16)') 'dg1d_fvwb mode=', trim(mode), ' n=', n, ' nx=', nx, ' steps=', iter, &
       ' t=', t
  if (len_trim(prefix) > 0) then
     open(newunit=unit, file=trim(prefix)//'.bin', access='stream', form='unformatted', status='replace')
     write(unit) int(nx, c_int32_t), int(n, c_int32_t), int(nvar, c_int32_t), int(iter, c_int32_t)
     write(unit) t, dt
     write(unit) delta_u, uinit
     close(unit)
  end if
  call fvwb_dg1d_dropin_close()

contains

  subroutine primitive(u, w)
    real(kind=8), intent(in) :: u(nvar)
    real(kind=8), intent(out) :: w(nvar)
    w(1) = u(1)
    w(2) = u(2) / w(1)
    w(3) = (gamma - 1.0) * (u(3) - 0.5 * w(1) * w(2)**2)
  end subroutine primitive

  subroutine conservative(w, u)
    real(kind=8), intent(in) :: w(nvar)
    real(kind=8), intent(out) :: u(nvar)
    u(1) = w(1)
    u(2) = w(1) * w(2)
    u(3) = w(3) / (gamma - 1.0) + 0.5 * w(1) * w(2)**2
  end subroutine conservative

  ! condinit cases 7 and 8 (the well-balanced test problems)
  subroutine initial_state(x, uu)
    real(kind=8), intent(in) :: x
    real(kind=8), intent(out) :: uu(nvar)
    real(kind=8) :: ww(nvar)
    select case (ninit)
    case (7)
       ww(1) = exp(-x)
       ww(2) = 0.
       ww(3) = exp(-x)
    case (8)
       ww(1) = exp(-x)
       ww(2) = 0.
       ww(3) = exp(-x) + pert * exp(-100 * (x - boxlen / 2.)**2)
    case default
       error stop 'dg1d_fvwb: ninit must be 7 or 8'
    end select
    call conservative(ww, uu)
  end subroutine initial_state

  real(kind=8) function speed(u)
    real(kind=8), intent(in) :: u(nvar)
    real(kind=8) :: w(nvar)
    call primitive(u, w)
    speed = abs(w(2)) + sqrt(gamma * max(w(3), 1d-10) / max(w(1), 1d-10))
  end function speed

  subroutine evolve_stages(t, dt, iter)
    real(kind=8), intent(out) :: t, dt
    integer, intent(out) :: iter
    real(kind=8), dimension(1:nvar, 1:n, 1:nx) :: dudt, w1, w2, w3, w4, nodes
    real(kind=8) :: cmax
    integer :: ivar, icell, i, intnode
    t = 0
    iter = 0
    dt = 0
    do while (t < tend)
       cmax = 0.0
       do icell = 1, nx
          cmax = max(cmax, speed(uinit(1:nvar, 1, icell)))
       end do
       dt = 0.9 * dx / cmax / (2.0 * dble(n) + 1.0)
       call compute_update_exact_delta(delta_u, u_eq, dudt)
       w1 = delta_u + 0.391752226571890 * dt * dudt
       call compute_update_exact_delta(w1, u_eq, dudt)
       w2 = 0.444370493651235 * delta_u + 0.555629506348765 * w1 + 0.368410593050371 * dt * dudt
       call compute_update_exact_delta(w2, u_eq, dudt)
       w3 = 0.620101851488403 * delta_u + 0.379898148511597 * w2 + 0.251891774271694 * dt * dudt
       call compute_update_exact_delta(w3, u_eq, dudt)
       w4 = 0.178079954393132 * delta_u + 0.821920045606868 * w3 + 0.544974750228521 * dt * dudt
       delta_u = 0.517231671970585 * w2 + 0.096059710526147 * w3 + 0.063692468666290 * dt * dudt
       call compute_update_exact_delta(w4, u_eq, dudt)
       delta_u = delta_u + 0.386708617503269 * w4 + 0.226007483236906 * dt * dudt
       ! reconstruction at legendre(x_quad), x_quad physical, as the reference does
       nodes(:, :, :) = 0.0
       do ivar = 1, nvar
          do icell = 1, nx
             do i = 1, n
                x_quad = (dble(icell) - 0.5) * dx + dx / 2.0 * chsi_quad(i)
                do intnode = 1, n
                   nodes(ivar, i, icell) = nodes(ivar, i, icell) + &
                        & delta_u(ivar, intnode, icell) * basis1d(x_quad, intnode - 1)
                end do
             end do
          end do
       end do
       uinit = u_eq + nodes
       t = t + dt
       iter = iter + 1
       write(*, *) 'time=', iter, t, dt
    end do
  end subroutine evolve_stages

  subroutine evolve_resident(t, dt, iter)
    real(kind=8), intent(out) :: t, dt
    integer, intent(out) :: iter
    type(c_ptr) :: h
    integer(c_int32_t) :: steps
    h = fvwb_dg1d_dropin_handle()
    call fvwb_dg1d_check(fvwb_dg1d_upload(h, delta_u, u_eq, uinit), h, 'fvwb_dg1d_upload')
    call fvwb_dg1d_check(fvwb_dg1d_advance(h, huge(steps), tend), h, 'fvwb_dg1d_advance')
    call fvwb_dg1d_check(fvwb_dg1d_download(h, delta_u, uinit), h, 'fvwb_dg1d_download')
    call fvwb_dg1d_check(fvwb_dg1d_get_time(h, t, dt, steps), h, 'fvwb_dg1d_get_time')
    iter = steps
  end subroutine evolve_resident

end program dg1d_fvwb
