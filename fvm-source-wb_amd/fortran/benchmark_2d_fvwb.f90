! benchmark_2d_fvwb — the 2D Euler-with-gravity benchmark run on MI355X through libfvwb.
!
! The driver of benchmark_2d.f90 (program main, :1-22): cell-centred coordinates
! on [0,boxlen_x]x[0,boxlen_y], the initial state selected by ninit, the
! primitive equilibrium selected by nequilibrium, the Heun time loop of evolve
! (:221-260: dt = 0.5*dx/cmax*cfl, w1 = u + dt L(u), u = u/2 + w1/2 + dt/2 L(w1),
! while t < tend, no clipping of the last step) and the ASCII snapshot format of
! output_file (:115-143: x, y, p - p_eq per cell as '(7(1PE12.5,1X))').
!
! usage: benchmark_2d_fvwb [stage|resident|resume] [tend] [prefix] [checkpoint]
!   stage     the reference's loop on the host; every residual and wave-speed
!             evaluation is the drop-in compute_update_exact / compute_max_speed
!             (state crosses PCIe twice per stage, as the reference's caller owns it)
!   resident  the state is uploaded once and advanced on the device to tend
!             (fvwb_advance), then downloaded; with a checkpoint path, the
!             device state at tend is also saved there (fvwb_checkpoint_save)
!   resume    the device state is loaded from the checkpoint path
!             (fvwb_checkpoint_load) and advanced from its time to tend
!   prefix    when given: <prefix>_init.dat and <prefix>_final.dat snapshots and
!             <prefix>.bin (int32 nx, ny, nvar, steps; real(8) t; u0, w_eq, u)
program benchmark_2d_fvwb
  use iso_c_binding
  use parameters_2d
  use fvwb_iface
  use fvwb_dropin_2d
  implicit none
  real(kind=8), allocatable :: x(:,:), y(:,:), u(:,:,:), u0(:,:,:), w_eq(:,:,:)
  character(len=512) :: arg, mode, prefix, ckpt
  integer :: nargs, iter, ios
  real(kind=8) :: t

  mode = 'stage'
  prefix = ''
  nargs = command_argument_count()
  if (nargs >= 1) call get_command_argument(1, mode)
  if (nargs >= 2) then
     call get_command_argument(2, arg)
     read(arg, *, iostat=ios) tend
     if (ios /= 0) error stop 'benchmark_2d_fvwb: tend must be a number'
  end if
  ckpt = ''
  if (nargs >= 3) call get_command_argument(3, prefix)
  if (nargs >= 4) call get_command_argument(4, ckpt)

  allocate(x(nx, ny), y(nx, ny), u(nvar, nx, ny), u0(nvar, nx, ny), w_eq(nvar, nx, ny))
  call cell_centres(x, y)
  call initial_state(x, y, u)
  call equilibrium_primitive(x, y, w_eq)
  u0 = u
  if (len_trim(prefix) > 0) call write_snapshot(x, y, u, w_eq, trim(prefix)//'_init.dat')

  select case (trim(mode))
  case ('stage')
     call evolve_stages(u, w_eq, t, iter)
  case ('resident', 'resume')
     call evolve_resident(u, w_eq, t, iter)
  case default
     error stop 'benchmark_2d_fvwb: mode must be stage, resident or resume'
  end select
  write(*, '(a,a,a,i0,a,i0,a,i0,a,1pe24.16)') 'benchmark_2d_fvwb mode=', trim(mode), ' nx=', nx, &
       ' ny=', ny, ' steps=', iter, ' t=', t

  if (len_trim(prefix) > 0) then
     call write_snapshot(x, y, u, w_eq, trim(prefix)//'_final.dat')
     call write_state(trim(prefix)//'.bin', u0, w_eq, u, iter, t)
  end if
  call fvwb_dropin_close()

contains

  subroutine cell_centres(x, y)
    real(kind=8), intent(out) :: x(nx, ny), y(nx, ny)
    real(kind=8) :: dx, dy
    integer :: i, j
    dx = boxlen_x / dble(nx)
    dy = boxlen_y / dble(ny)
    do j = 1, ny
       do i = 1, nx
          x(i, j) = (dble(i) - 0.5d0) * dx
          y(i, j) = (dble(j) - 0.5d0) * dy
       end do
    end do
  end subroutine cell_centres

  ! isothermal atmosphere rho = rho0 e^{-k s}, p = p0 e^{-k s}, s = x + y,
  ! k = rho0 g / p0; rho0 = 1.21 is a default-real (FP32) constant as in the
  ! reference set-up (benchmark_2d.f90:195-202)
  subroutine atmosphere(x, y, rho0, w)
    real(kind=8), intent(in) :: x(nx, ny), y(nx, ny), rho0
    real(kind=8), intent(out) :: w(nvar, nx, ny)
    real(kind=8), parameter :: p0 = 1.0d0, g = 1.0d0
    w(1, :, :) = rho0 * exp(-(rho0 * g / p0) * (x + y))
    w(2, :, :) = 0.0d0
    w(3, :, :) = 0.0d0
    w(4, :, :) = p0 * exp(-(rho0 * g / p0) * (x + y))
  end subroutine atmosphere

  subroutine equilibrium_primitive(x, y, w)
    real(kind=8), intent(in) :: x(nx, ny), y(nx, ny)
    real(kind=8), intent(out) :: w(nvar, nx, ny)
    real(kind=8) :: rho0
    select case (nequilibrium)
    case (1)
       call atmosphere(x, y, 1.0d0, w)
    case (2, 3)
       rho0 = 1.21
       call atmosphere(x, y, rho0, w)
    case default
       w = 0.0d0
    end select
  end subroutine equilibrium_primitive

  subroutine initial_state(x, y, u)
    real(kind=8), intent(in) :: x(nx, ny), y(nx, ny)
    real(kind=8), intent(out) :: u(nvar, nx, ny)
    real(kind=8) :: w(nvar, nx, ny), rho0, k, c03
    integer :: i, j
    rho0 = 1.21
    c03 = 0.3
    k = rho0 * 1.0d0 / 1.0d0
    select case (ninit)
    case (1)
       call atmosphere(x, y, 1.0d0, w)
    case (2)
       call atmosphere(x, y, rho0, w)
    case (3)
       ! atmosphere plus a Gaussian pressure pulse of amplitude eta centred at
       ! (0.3, 0.3); the centre is a default-real constant as in the reference
       call atmosphere(x, y, rho0, w)
       w(4, :, :) = w(4, :, :) + eta * exp(-100.0d0 * k * ((x - c03)**2 + (y - c03)**2))
    case (4)
       ! four-quadrant 2D Riemann problem (benchmark_2d.f90:103-131); the states
       ! are default-real constants, widened to real(kind=8)
       do j = 1, ny
          do i = 1, nx
             if (x(i, j) >= 0.5d0 .and. y(i, j) >= 0.5d0) then
                w(:, i, j) = [1.5, 0.0, 0.0, 1.5]
             else if (x(i, j) < 0.5d0 .and. y(i, j) >= 0.5d0) then
                w(:, i, j) = [0.5323, 1.206, 0.0, 0.3]
             else if (x(i, j) < 0.5d0 .and. y(i, j) < 0.5d0) then
                w(:, i, j) = [0.138, 1.206, 1.206, 0.029]
             else
                w(:, i, j) = [0.5323, 0.0, 1.206, 0.3]
             end if
          end do
       end do
    case default
       error stop 'benchmark_2d_fvwb: unknown ninit'
    end select
    u(1, :, :) = w(1, :, :)
    u(2, :, :) = w(1, :, :) * w(2, :, :)
    u(3, :, :) = w(1, :, :) * w(3, :, :)
    u(4, :, :) = w(4, :, :) / (gamma - 1.0d0) + 0.5d0 * (w(1, :, :) * (w(2, :, :)**2 + w(3, :, :)**2))
  end subroutine initial_state

  ! The reference's time loop with the per-stage drop-ins.
  subroutine evolve_stages(u, w_eq, t, iter)
    real(kind=8), intent(inout) :: u(nvar, nx, ny)
    real(kind=8), intent(in) :: w_eq(nvar, nx, ny)
    real(kind=8), intent(out) :: t
    integer, intent(out) :: iter
    real(kind=8), allocatable :: dudt(:,:,:), w1(:,:,:)
    real(kind=8) :: dt, cmax, dx
    allocate(dudt(nvar, nx, ny), w1(nvar, nx, ny))
    dx = boxlen_x / dble(nx)
    t = 0.0d0
    iter = 0
    do while (t < tend)
       call compute_max_speed(u, cmax)
       dt = 0.5d0 * dx / cmax * cfl
       call compute_update_exact(u, w_eq, dudt)
       w1 = u + dt * dudt
       call compute_update_exact(w1, w_eq, dudt)
       u = 0.5d0 * u + 0.5d0 * w1 + 0.5d0 * dt * dudt
       t = t + dt
       iter = iter + 1
    end do
  end subroutine evolve_stages

  ! The same time loop with the state resident in HBM.
  subroutine evolve_resident(u, w_eq, t, iter)
    real(kind=8), intent(inout) :: u(nvar, nx, ny)
    real(kind=8), intent(in) :: w_eq(nvar, nx, ny)
    real(kind=8), intent(out) :: t
    integer, intent(out) :: iter
    type(c_ptr) :: h
    integer(c_int32_t) :: steps
    h = fvwb_dropin_handle()
    if (trim(mode) == 'resume') then
       if (len_trim(ckpt) == 0) error stop 'benchmark_2d_fvwb: resume needs a checkpoint path'
       call fvwb_check(fvwb_checkpoint_load(h, trim(ckpt)//c_null_char), h, 'fvwb_checkpoint_load')
    else
       call fvwb_check(fvwb_upload(h, u, w_eq), h, 'fvwb_upload')
    end if
    call fvwb_check(fvwb_advance(h, huge(steps), tend, t, steps), h, 'fvwb_advance')
    if (trim(mode) == 'resident' .and. len_trim(ckpt) > 0) &
         call fvwb_check(fvwb_checkpoint_save(h, trim(ckpt)//c_null_char), h, 'fvwb_checkpoint_save')
    call fvwb_check(fvwb_download(h, u), h, 'fvwb_download')
    iter = steps
  end subroutine evolve_resident

  ! x, y, p - p_eq per cell, x outer and y inner, as the reference's snapshots.
  subroutine write_snapshot(x, y, u, w_eq, path)
    real(kind=8), intent(in) :: x(nx, ny), y(nx, ny), u(nvar, nx, ny), w_eq(nvar, nx, ny)
    character(len=*), intent(in) :: path
    integer :: i, j, unit
    real(kind=8) :: p
    open(newunit=unit, file=path, form='formatted', status='replace')
    do i = 1, nx
       do j = 1, ny
          p = (gamma - 1.0d0) * (u(4, i, j) - 0.5d0 * u(1, i, j) * &
               ((u(2, i, j) / u(1, i, j))**2 + (u(3, i, j) / u(1, i, j))**2))
          write(unit, '(7(1PE12.5,1X))') x(i, j), y(i, j), p - w_eq(4, i, j)
       end do
    end do
    close(unit)
  end subroutine write_snapshot

  subroutine write_state(path, u0, w_eq, u, iter, t)
    character(len=*), intent(in) :: path
    real(kind=8), intent(in) :: u0(nvar, nx, ny), w_eq(nvar, nx, ny), u(nvar, nx, ny), t
    integer, intent(in) :: iter
    integer :: unit
    open(newunit=unit, file=path, access='stream', form='unformatted', status='replace')
    write(unit) int(nx, c_int32_t), int(ny, c_int32_t), int(nvar, c_int32_t), int(iter, c_int32_t)
    write(unit) t
    write(unit) u0, w_eq, u
    close(unit)
  end subroutine write_state

end program benchmark_2d_fvwb
