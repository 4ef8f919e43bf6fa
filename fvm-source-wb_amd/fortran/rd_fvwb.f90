! rd_fvwb — the residual-distribution program benchmark_rd.f90 (node grid of
! get_coords, get_initial_conditions cases 1, 2 and 8, evolve with solver 'RD1')
! over the GPU drop-ins.
!
!   rd_fvwb stage    [tend] [prefix]   evolve's loop on the host, compute_max_speed
!                                      and compute_update_rd on the GPU
!   rd_fvwb resident [tend] [prefix] [checkpoint] [nsteps]
!                                      the whole loop resident on the device, at most
!                                      nsteps steps; with a checkpoint path the device
!                                      state at the end is also saved there
!                                      (fvwb_rd_checkpoint_save)
!   rd_fvwb resume   [tend] [prefix] checkpoint
!                                      the device state and clock are loaded from the
!                                      checkpoint (fvwb_rd_checkpoint_load) and
!                                      advanced from their time to tend
!
! Writes prefix.bin (stream): int32 nx, ny, steps, 0; real(8) t; the initial
! state u0(4,nx,ny); the final state u(4,nx,ny).
program rd_fvwb
  use parameters_rd
  use iso_c_binding
  use fvwb_iface
  use fvwb_dropin_rd
  implicit none
  real(kind=8), dimension(1:nx, 1:ny, 1:mx, 1:my) :: x, y
  real(kind=8), dimension(1:nvar, 1:nx, 1:ny, 1:mx, 1:my) :: u, u0, diff, dudt, up
  real(kind=8) :: t, dt, cmax, dx, dy, t_end
  integer :: iter, i, j, nargs, unit
  integer(c_int32_t) :: max_steps
  integer(c_int64_t) :: steps
  character(len=32) :: mode, arg
  character(len=256) :: prefix, ckpt
  type(c_ptr) :: h
  external :: compute_update_rd, compute_max_speed

  mode = 'stage'
  t_end = tend
  prefix = 'rd'
  nargs = command_argument_count()
  if (nargs >= 1) call get_command_argument(1, mode)
  if (nargs >= 2) then
     call get_command_argument(2, arg)
     read(arg, *) t_end
  end if
  if (nargs >= 3) call get_command_argument(3, prefix)
  ckpt = ''
  if (nargs >= 4) call get_command_argument(4, ckpt)
  max_steps = huge(1_c_int32_t)
  if (nargs >= 5) then
     call get_command_argument(5, arg)
     read(arg, *) max_steps
  end if

  ! get_coords (benchmark_rd.f90:24-51): nodes at (i-1) dx, (j-1) dy
  dx = boxlen_x / dble(nx - 1)
  dy = boxlen_y / dble(ny - 1)
  do i = 1, nx
     do j = 1, ny
        x(i, j, 1, 1) = (i - 1) * dx
        y(i, j, 1, 1) = (j - 1) * dy
     end do
  end do
  call initial_state(x, y, u)
  u0 = u

  t = 0
  iter = 0
  if (trim(mode) == 'stage') then
     ! evolve (benchmark_rd.f90:356-385), solver 'RD1'
     do while (t < t_end)
        call compute_max_speed(u, cmax)
        dt = min(cfl * dy * dx / cmax / (2.0 * dble(1) + 1.0), t_end - t)
        diff = 0
        call compute_update_rd(u, diff, dt, x, y, dudt)
        up = u - dudt / (dx * dy)
        diff = up - u
        call compute_update_rd(u, diff, dt, x, y, dudt)
        u = up - dudt / (dx * dy)
        t = t + dt
        iter = iter + 1
        write(*, *) 'time=', iter, t, dt, cmax
     end do
     steps = iter
     call fvwb_rd_dropin_close()
  else if (trim(mode) == 'resident' .or. trim(mode) == 'resume') then
     call fvwb_rd_check(fvwb_rd_create(h, fvwb_rd_dropin_params()), c_null_ptr, 'fvwb_rd_create')
     if (trim(mode) == 'resume') then
        if (len_trim(ckpt) == 0) error stop 'rd_fvwb: resume needs a checkpoint path'
        call fvwb_rd_check(fvwb_rd_checkpoint_load(h, trim(ckpt)//c_null_char), h, 'fvwb_rd_checkpoint_load')
     else
        call fvwb_rd_check(fvwb_rd_upload(h, u), h, 'fvwb_rd_upload')
     end if
     call fvwb_rd_check(fvwb_rd_advance(h, max_steps, t_end), h, 'fvwb_rd_advance')
     if (trim(mode) == 'resident' .and. len_trim(ckpt) > 0) &
          call fvwb_rd_check(fvwb_rd_checkpoint_save(h, trim(ckpt)//c_null_char), h, 'fvwb_rd_checkpoint_save')
     call fvwb_rd_check(fvwb_rd_download(h, u), h, 'fvwb_rd_download')
     call fvwb_rd_check(fvwb_rd_get_time(h, t, dt, steps), h, 'fvwb_rd_get_time')
     call fvwb_rd_check(fvwb_rd_destroy(h), c_null_ptr, 'fvwb_rd_destroy')
     write(*, *) 'time=', steps, t, dt
  else
     write(0, *) 'usage: rd_fvwb stage|resident|resume [tend] [prefix] [checkpoint] [nsteps]'
     error stop 2
  end if

  open(newunit=unit, file=trim(prefix)//'.bin', access='stream', form='unformatted', status='replace')
  write(unit) int(nx, c_int32_t), int(ny, c_int32_t), int(steps, c_int32_t), 0_c_int32_t, t
  write(unit) u0, u
  close(unit)

contains

  ! get_initial_conditions (benchmark_rd.f90:53-265), cases 1, 2 and 8, then
  ! compute_conservative (:535-547); default-real literals as the reference writes them
  subroutine initial_state(x, y, u)
    real(kind=8), dimension(1:nx, 1:ny, 1:mx, 1:my), intent(in) :: x, y
    real(kind=8), dimension(1:nvar, 1:nx, 1:ny, 1:mx, 1:my), intent(out) :: u
    real(kind=8), dimension(1:nvar, 1:nx, 1:ny, 1:mx, 1:my) :: w
    real(kind=8) :: rho_0, p_0, g
    select case (ninit)
    case (1)
       w(1, :, :, :, :) = 1 + exp(-((x - boxlen_x / 2.)**2 + (y - boxlen_y / 2.)**2) * 20)
       w(2, :, :, :, :) = 1.
       w(3, :, :, :, :) = 1.
       w(4, :, :, :, :) = 1.
    case (2)
       rho_0 = 1.21
       p_0 = 1
       g = 1
       w(1, :, :, :, :) = rho_0 * exp(-(rho_0 * g / p_0) * (x + y))
       w(2, :, :, :, :) = 0
       w(3, :, :, :, :) = 0
       w(4, :, :, :, :) = p_0 * exp(-(rho_0 * g / p_0) * (x + y))
    case (8)
       w(1, :, :, :, :) = 1.21 * exp(-1.21 * (x + y))
       w(2, :, :, :, :) = 0
       w(3, :, :, :, :) = 0
       w(4, :, :, :, :) = exp(-1.21 * (x + y)) + eta * exp(-100 * 1.21 * ((x - 0.3)**2 + (y - 0.3)**2))
    case default
       write(0, *) 'rd_fvwb: ninit must be 1, 2 or 8'
       error stop 2
    end select
    u(1, :, :, :, :) = w(1, :, :, :, :)
    u(2, :, :, :, :) = w(1, :, :, :, :) * w(2, :, :, :, :)
    u(3, :, :, :, :) = w(1, :, :, :, :) * w(3, :, :, :, :)
    u(4, :, :, :, :) = w(4, :, :, :, :) / (gamma - 1.) + 0.5 * (w(1, :, :, :, :) * (w(2, :, :, :, :)**2 &
         + w(3, :, :, :, :)**2))
  end subroutine initial_state

end program rd_fvwb
