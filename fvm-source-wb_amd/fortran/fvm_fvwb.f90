! fvm_fvwb — the stand-alone 1D program fvm.f90 run on MI355X through libfvwb.
!
! The program of fvm.f90:1-98: cell centres (i-0.5)*dx, the initial state of
! condinit (:100-175, ninit 1-7), the RK2 loop (:53-72) with
! dt = 0.8*dx/cmax/(2n+1), every residual and wave-speed evaluation through the
! drop-in compute_update / compute_max_speed, and the snapshots value00000.dat
! and value99999.dat (x, rho, v, p as '(7(1PE12.5,1X))').
!
! usage: fvm_fvwb [tend] [prefix]
!   prefix  the snapshots become <prefix>value00000.dat and <prefix>value99999.dat,
!           and <prefix>.bin holds int32 nx, nvar, steps; real(8) t; u0, u
program fvm_fvwb
  use iso_c_binding
  use fvm_commons
  use fvwb_dropin_fvm
  implicit none
  real(kind=8), dimension(1:nvar, 1:nx) :: u, u0, dudt, w1
  real(kind=8), dimension(1:nvar) :: uu
  character(len=512) :: arg, prefix
  integer :: nargs, ios, icell, iter, unit
  real(kind=8) :: dx, xcell, t, dt, cmax

  prefix = ''
  nargs = command_argument_count()
  if (nargs >= 1) then
     call get_command_argument(1, arg)
     read(arg, *, iostat=ios) tend
     if (ios /= 0) error stop 'fvm_fvwb: tend must be a number'
  end if
  if (nargs >= 2) call get_command_argument(2, prefix)

  dx = boxlen / dble(nx)
  do icell = 1, nx
     xcell = (dble(icell) - 0.5) * dx
     call initial_state(xcell, uu)
     u(1:nvar, icell) = uu
  end do
  u0 = u
  call write_primitive(trim(prefix)//'value00000.dat', u)

  t = 0
  iter = 0
  do while (t < tend)
     call compute_max_speed(u, cmax)
     dt = 0.8 * dx / cmax / (2.0 * dble(n) + 1.0)
     call compute_update(u, dudt)
     w1 = u + dt * dudt
     call compute_update(w1, dudt)
     u = 0.5 * u + 0.5 * w1 + 0.5 * dt * dudt
     t = t + dt
     iter = iter + 1
  end do

  call write_primitive(trim(prefix)//'value99999.dat', u)
  write(*, '(a,i0,a,i0,a,i0,a,i0,a,1pe24.16)') 'fvm_fvwb nx=', nx, ' bc=', bc, ' source=', source, &
       ' steps=', iter, ' t=', t
  if (len_trim(prefix) > 0) then
     open(newunit=unit, file=trim(prefix)//'.bin', access='stream', form='unformatted', status='replace')
     write(unit) int(nx, c_int32_t), int(nvar, c_int32_t), int(iter, c_int32_t)
     write(unit) t
     write(unit) u0, u
     close(unit)
  end if
  call fvwb_fvm_dropin_close()

contains

  ! the initial conditions of condinit (fvm.f90:100-175), literals as written there
  subroutine initial_state(x, uu)
    real(kind=8), intent(in) :: x
    real(kind=8), intent(out) :: uu(nvar)
    real(kind=8) :: ww(nvar)
    real(kind=8) :: dpi
    dpi = acos(-1d0)
    select case (ninit)
    case (1)
       ww = [1.0 + 0.5 * sin(2.0 * dpi * x), 1.0d0, 1.0d0]
    case (2)
       ww(1) = merge(2.0d0, 1.0d0, abs(x - 0.5) < 0.25)
       ww(2:3) = 1.0
    case (3)
       ww(1) = 1. + exp(-(x - 0.25)**2 / 2.0 / 0.05**2)
       if (abs(x - 0.7) < 0.1) ww(1) = ww(1) + 1.
       ww(2:3) = 1.0
    case (4)
       if (abs(x - 0.25) < 0.25) then
          ww = [1.0, 0.0, 1.0]
       else
          ww = [0.125, 0.0, 0.1]
       end if
    case (5)
       if (x < 0.1) then
          ww = [1.0, 0.0, 1000.0]
       else if (x < 0.9) then
          ww = [1.0, 0.0, 0.01]
       else
          ww = [1.0, 0.0, 100.]
       end if
    case (6)
       if (x < 10.0) then
          ww = [3.857143, -0.920279, 10.333333]
       else
          ww(1) = 1.0 + 0.2 * sin(5.0 * (x - 10.0))
          ww(2) = -3.549648
          ww(3) = 1.0
       end if
    case (7)
       ww(1) = (1 - (gamma - 1) / (dble(gamma)) * x)
       ww(2) = 0
       ww(3) = ww(1)**gamma
    case default
       error stop 'fvm_fvwb: unknown ninit'
    end select
    uu(1) = ww(1)
    uu(2) = ww(1) * ww(2)
    uu(3) = ww(3) / (gamma - 1.0) + 0.5 * ww(1) * ww(2)**2
  end subroutine initial_state

  subroutine write_primitive(path, u)
    character(len=*), intent(in) :: path
    real(kind=8), intent(in) :: u(nvar, nx)
    integer :: k, unit
    real(kind=8) :: w(nvar)
    open(newunit=unit, file=path, form='formatted', status='replace')
    do k = 1, nx
       w(1) = u(1, k)
       w(2) = u(2, k) / w(1)
       w(3) = (gamma - 1.0) * (u(3, k) - 0.5 * w(1) * w(2)**2)
       write(unit, '(7(1PE12.5,1X))') (dble(k) - 0.5) * dx, w
    end do
    close(unit)
  end subroutine write_primitive

end program fvm_fvwb
