! benchmark_2d_dg_fvwb — the reference's 2D DG program (2d/benchmark_2d_dg.f90:
! main, get_coords, get_initial_conditions cases 1 and 2,
! get_equilibrium_solution, output_file, evolve) over the GPU drop-ins of
! fvwb_dropin_dg2d.f90.
!
!   benchmark_2d_dg_fvwb stage    [tend] [prefix]  evolve's loop on the host; compute_update,
!                                                  compute_max_speed, apply_limiter and the
!                                                  projections on the GPU
!   benchmark_2d_dg_fvwb resident [tend] [prefix]  the loop resident on the device
!                                                  (fvwb_dgq_advance), snapshots between
!
! Like the reference it writes its snapshots into longrotation2_o1/ (which must
! exist): initcond.dat, derp.dat (the projected, then the limited, state),
! SIMnnnnn.dat every `interval` steps when make_movie is set (:759-766),
! fintwo.dat, all in '(7(1PE12.5,1X))'; it prints 'time=' per step in stage
! mode and compute_error's norms at the end.  It also writes prefix.bin
! (stream): int32 nx, ny, mx, steps; real(8) t; the final nodal state
! u(1:4,1:nx,1:ny,1:mx,1:my).
program benchmark_2d_dg_fvwb
  use parameters_dg_2d
  use iso_c_binding
  use fvwb_dropin_dg2d
  implicit none
  real(kind=8), dimension(1:nx, 1:ny, 1:mx, 1:my) :: x, y
  real(kind=8), dimension(1:nvar, 1:nx, 1:ny, 1:mx, 1:my) :: u, u_eq
  character(len=32) :: mode, arg
  character(len=256) :: prefix
  integer :: nargs, unit, iters
  real(kind=8) :: t_final

  mode = 'stage'
  prefix = 'dg2d'
  nargs = command_argument_count()
  if (nargs >= 1) call get_command_argument(1, mode)
  if (nargs >= 2) then
     call get_command_argument(2, arg)
     read(arg, *) tend
  end if
  if (nargs >= 3) call get_command_argument(3, prefix)
  if (trim(mode) /= 'stage' .and. trim(mode) /= 'resident') then
     write(0, *) 'usage: benchmark_2d_dg_fvwb stage|resident [tend] [prefix]'
     error stop 2
  end if

  call get_coords(x, y, nx, ny, mx, my)
  call get_initial_conditions(x, y, u, nx, ny, mx, my)
  call output_file(x, y, u, 1, 'initcond')
  call get_equilibrium_solution(x, y, u_eq, nx, ny, mx, my)
  call evolve(u, x, y, u_eq, trim(mode) == 'resident', iters, t_final)
  call output_file(x, y, u, 1, 'fintwo')
  call fvwb_dgq_dropin_close()

  open(newunit=unit, file=trim(prefix)//'.bin', access='stream', form='unformatted', status='replace')
  write(unit) int(nx, c_int32_t), int(ny, c_int32_t), int(mx, c_int32_t), int(iters, c_int32_t), t_final
  write(unit) u
  close(unit)
end program benchmark_2d_dg_fvwb

! get_coords (:93-120): the Gauss points of each cell
subroutine get_coords(x, y, size_x, size_y, order_x, order_y)
  use parameters_dg_2d
  implicit none
  integer :: size_x, size_y, order_x, order_y
  real(kind=8), dimension(1:size_x, 1:size_y, 1:order_x, 1:order_y) :: x, y
  real(kind=8) :: dx, dy
  integer :: i, j, a, b
  external :: gl_quadrature
  call gl_quadrature(x_quad, w_x_quad, order_x)
  call gl_quadrature(y_quad, w_y_quad, order_y)
  dx = boxlen_x / dble(size_x)
  dy = boxlen_y / dble(size_y)
  do i = 1, size_x
     do j = 1, size_y
        do a = 1, order_x
           do b = 1, order_y
              x(i, j, a, b) = (i - 0.5) * dx + dx / 2.0 * x_quad(a)
              y(i, j, a, b) = (j - 0.5) * dy + dy / 2.0 * y_quad(b)
           end do
        end do
     end do
  end do
end subroutine get_coords

! compute_primitive (:891-902) and compute_conservative (:905-917)
subroutine compute_primitive(u, w, size_x, size_y, order_x, order_y)
  use parameters_dg_2d
  implicit none
  integer :: size_x, size_y, order_x, order_y
  real(kind=8), dimension(1:nvar, 1:size_x, 1:size_y, 1:order_x, 1:order_y) :: u, w
  w(1, :, :, :, :) = max(u(1, :, :, :, :), 10e-10)
  w(2, :, :, :, :) = u(2, :, :, :, :) / w(1, :, :, :, :)
  w(3, :, :, :, :) = u(3, :, :, :, :) / w(1, :, :, :, :)
  w(4, :, :, :, :) = (gamma - 1.0) * (u(4, :, :, :, :) - 0.5 * w(1, :, :, :, :) * (w(2, :, :, :, :)**2 &
       + w(3, :, :, :, :)**2))
end subroutine compute_primitive

subroutine compute_conservative(ww, u, size_x, size_y, order_x, order_y)
  use parameters_dg_2d
  implicit none
  integer :: size_x, size_y, order_x, order_y
  real(kind=8), dimension(1:nvar, 1:size_x, 1:size_y, 1:order_x, 1:order_y) :: ww, u
  u(1, :, :, :, :) = ww(1, :, :, :, :)
  u(2, :, :, :, :) = ww(1, :, :, :, :) * ww(2, :, :, :, :)
  u(3, :, :, :, :) = ww(1, :, :, :, :) * ww(3, :, :, :, :)
  u(4, :, :, :, :) = ww(4, :, :, :, :) / (gamma - 1.) + 0.5 * (ww(1, :, :, :, :) * (ww(2, :, :, :, :)**2 &
       + ww(3, :, :, :, :)**2))
end subroutine compute_conservative

! get_initial_conditions (:121-466), cases 1 (a pulse advected diagonally)
! and 2 (the rho_0 = 1.21 atmosphere with a pressure pulse); default-real
! literals as the reference writes them, since test problems are data
subroutine get_initial_conditions(x, y, u, size_x, size_y, order_x, order_y)
  use parameters_dg_2d
  implicit none
  integer :: size_x, size_y, order_x, order_y
  real(kind=8), dimension(1:nvar, 1:size_x, 1:size_y, 1:order_x, 1:order_y) :: u, w
  real(kind=8), dimension(1:size_x, 1:size_y, 1:order_x, 1:order_y) :: x, y
  real(kind=8) :: rho_0, p_0, g
  select case (ninit)
  case (1)
     w(1, :, :, :, :) = exp(-((x - boxlen_x / 2.)**2 + (y - boxlen_y / 2.)**2) * 10)
     w(2, :, :, :, :) = 1.0
     w(3, :, :, :, :) = 1.0
     w(4, :, :, :, :) = minval(w(1, :, :, :, :))
  case (2)
     rho_0 = 1.21
     p_0 = 1.
     g = 1.
     w(1, :, :, :, :) = rho_0 * exp(-(rho_0 * g / p_0) * (x + y))
     w(2, :, :, :, :) = 0
     w(3, :, :, :, :) = 0
     w(4, :, :, :, :) = p_0 * exp(-(rho_0 * g / p_0) * (x + y)) + eta * &
          exp(-100 * (rho_0 * g / p_0) * ((x - 0.3)**2 + (y - 0.3)**2))
  case default
     write(0, *) 'benchmark_2d_dg_fvwb: ninit must be 1 or 2'
     error stop 2
  end select
  call compute_conservative(w, u, size_x, size_y, order_x, order_y)
end subroutine get_initial_conditions

! get_equilibrium_solution (:597-622), primitive
subroutine get_equilibrium_solution(x, y, w, size_x, size_y, order_x, order_y)
  use parameters_dg_2d
  implicit none
  integer :: size_x, size_y, order_x, order_y
  real(kind=8), dimension(1:nvar, 1:size_x, 1:size_y, 1:order_x, 1:order_y) :: w
  real(kind=8), dimension(1:size_x, 1:size_y, 1:order_x, 1:order_y) :: x, y
  real(kind=8) :: rho_0, p_0, g
  select case (nequilibrium)
  case (1)
     w(1, :, :, :, :) = exp(-(x + y))
     w(2, :, :, :, :) = 0
     w(3, :, :, :, :) = 0
     w(4, :, :, :, :) = exp(-(x + y))
  case (2)
     rho_0 = 1.21
     p_0 = 1
     g = 1
     w(1, :, :, :, :) = rho_0 * exp(-(rho_0 * g / p_0) * (x + y))
     w(2, :, :, :, :) = 0
     w(3, :, :, :, :) = 0
     w(4, :, :, :, :) = p_0 * exp(-(rho_0 * g / p_0) * (x + y))
  case default
     w(:, :, :, :, :) = 0.0
  end select
end subroutine get_equilibrium_solution

! output_file (:468-495): per cell (x outer, y inner) the first node's
! coordinates and primitive state minus the equilibrium there
subroutine output_file(x_values, y_values, function_values, var, filen)
  use parameters_dg_2d
  implicit none
  real(kind=8), dimension(1:nvar, 1:nx, 1:ny, 1:mx, 1:my) :: function_values
  real(kind=8), dimension(1:nx, 1:ny, 1:mx, 1:my) :: x_values, y_values
  integer :: var
  character(len=*) :: filen
  real(kind=8), dimension(1:nvar) :: w, w_eq
  real(kind=8) :: xcell, ycell
  integer :: icell, jcell, ivar, unit
  open(newunit=unit, status='replace', file='longrotation2_o1/'//trim(filen)//'.dat', form='formatted')
  do icell = 1, nx
     do jcell = 1, ny
        xcell = x_values(icell, jcell, 1, 1)
        ycell = y_values(icell, jcell, 1, 1)
        call get_equilibrium_solution([xcell], [ycell], w_eq, 1, 1, 1, 1)
        call compute_primitive(function_values(1:nvar, icell, jcell, 1, 1), w, 1, 1, 1, 1)
        write(unit, '(7(1PE12.5,1X))') xcell, ycell, (w(ivar) - w_eq(ivar), ivar = var, nvar)
     end do
  end do
  close(unit)
end subroutine output_file

! evolve (:624-775): project, limit, then while t < tend the dt of :671 and the
! solver's stages with apply_limiter after each; the snapshot cadence of
! :759-766; compute_error at the end; u returns the final nodal state.
subroutine evolve(u, x, y, u_eq, resident, iters, t_final)
  use parameters_dg_2d
  use iso_c_binding
  use fvwb_iface
  use fvwb_dropin_dg2d
  implicit none
  real(kind=8), dimension(1:nvar, 1:nx, 1:ny, 1:mx, 1:my) :: u, u_eq, u_anal
  real(kind=8), dimension(1:nx, 1:ny, 1:mx, 1:my) :: x, y
  logical :: resident
  integer :: iters
  real(kind=8) :: t_final
  real(kind=8), dimension(1:nvar, 1:nx, 1:ny, 1:mx, 1:my) :: dudt, w1, w2, w3, w4, w5, delta_u, nodes
  real(kind=8) :: t, dt, gll_w_1, cmax, dx, cs_max, v_xmax, v_ymax, dt_last
  integer :: iter, snap_counter
  integer(c_int64_t) :: steps
  character(len=10) :: filename
  type(c_ptr) :: h
  external :: compute_update, compute_max_speed, apply_limiter, get_modes_from_nodes, get_nodes_from_modes
  external :: compute_error, output_file

  dx = boxlen_x / dble(nx)
  delta_u = 0.0
  call get_modes_from_nodes(u, delta_u, nx, ny, mx, my)
  call get_nodes_from_modes(delta_u, nodes, nx, ny, mx, my)
  call output_file(x, y, nodes, 1, 'derp')
  if (mx == 1) then
     gll_w_1 = 1.
  else
     gll_w_1 = 1. / dble(gll * (gll - 1) + 1e-10)
  end if
  t = 0
  iter = 0
  snap_counter = 0
  call apply_limiter(delta_u)
  call get_nodes_from_modes(delta_u, nodes, nx, ny, mx, my)
  call output_file(x, y, nodes, 1, 'derp')

  if (resident) then
     h = fvwb_dgq_dropin_handle()
     call fvwb_dgq_check(fvwb_dgq_upload(h, delta_u, x, y), h, 'fvwb_dgq_upload')
     do while (t < tend)
        ! up to the next snapshot step, or to tend
        call fvwb_dgq_check(fvwb_dgq_advance(h, int(interval - modulo(iter, interval), c_int32_t), tend), h, &
             'fvwb_dgq_advance')
        call fvwb_dgq_check(fvwb_dgq_get_time(h, t, dt_last, steps), h, 'fvwb_dgq_get_time')
        iter = int(steps)
        if (make_movie .and. modulo(iter, interval) == 0) then
           call fvwb_dgq_check(fvwb_dgq_download(h, delta_u), h, 'fvwb_dgq_download')
           call get_nodes_from_modes(delta_u, nodes, nx, ny, mx, my)
           snap_counter = snap_counter + 1
           write(filename, '(a, i5.5)') 'SIM', snap_counter
           call output_file(x, y, nodes, 1, filename)
        end if
     end do
     call fvwb_dgq_check(fvwb_dgq_download(h, delta_u), h, 'fvwb_dgq_download')
  else
     do while (t < tend)
        call compute_max_speed(delta_u(:, :, :, 1, 1), cs_max, v_xmax, v_ymax, cmax)
        dt = min(tend - t, cfl * min(1. / dble(2 * 4 + 1), gll_w_1 / 2.) / ((abs(v_xmax) + (cs_max)) / dx + &
             (abs(v_ymax) + (cs_max)) / dx))
        select case (solver)
        case ('EQL')
           call compute_update(delta_u, x, y, u_eq, dudt)
           w1 = delta_u + dt * dudt
           call apply_limiter(w1)
           call compute_update(w1, x, y, u_eq, dudt)
           delta_u = 0.5 * delta_u + 0.5 * w1 + 0.5 * dt * dudt
           call apply_limiter(delta_u)
        case ('RK4')
           call compute_update(delta_u, x, y, u_eq, dudt)
           w1 = delta_u + 0.391752226571890 * dt * dudt
           call apply_limiter(w1)
           call compute_update(w1, x, y, u_eq, dudt)
           w2 = 0.444370493651235 * delta_u + 0.555629506348765 * w1 + 0.368410593050371 * dt * dudt
           call apply_limiter(w2)
           call compute_update(w2, x, y, u_eq, dudt)
           w3 = 0.620101851488403 * delta_u + 0.379898148511597 * w2 + 0.251891774271694 * dt * dudt
           call apply_limiter(w3)
           call compute_update(w3, x, y, u_eq, dudt)
           w4 = 0.178079954393132 * delta_u + 0.821920045606868 * w3 + 0.544974750228521 * dt * dudt
           call apply_limiter(w4)
           call compute_update(w3, x, y, u_eq, dudt)
           w5 = 0.00683325884039 * delta_u + 0.51723167208978 * w2 + 0.12759831133288 * w3 + &
                0.34833675773694 * w4 + 0.08460416338212 * dt * dudt
           call compute_update(w4, x, y, u_eq, dudt)
           delta_u = w5 + 0.22600748319395 * dt * dudt
           call apply_limiter(delta_u)
        case ('SS4')
           call compute_update(delta_u, x, y, u_eq, dudt)
           w1 = delta_u + 0.39175222700392 * dt * dudt
           call apply_limiter(w1)
           call compute_update(w1, x, y, u_eq, dudt)
           w2 = 0.44437049406734 * delta_u + 0.55562950593266 * w1 + 0.36841059262959 * dt * dudt
           call apply_limiter(w2)
           call compute_update(w2, x, y, u_eq, dudt)
           w3 = 0.6201018513854 * delta_u + 0.37989814861460 * w2 + 0.25189177424738 * dt * dudt
           call apply_limiter(w3)
           call compute_update(w3, x, y, u_eq, dudt)
           w4 = 0.17807995410773 * delta_u + 0.82192004589227 * w3 + 0.54497475021237 * dt * dudt
           call apply_limiter(w4)
           call compute_update(w3, x, y, u_eq, dudt)
           w5 = 0.00683325884039 * delta_u + 0.51723167208978 * w2 + 0.12759831133288 * w3 + &
                0.34833675773694 * w4 + 0.08460416338212 * dt * dudt
           call compute_update(w4, x, y, u_eq, dudt)
           delta_u = w5 + 0.22600748319395 * dt * dudt
           call apply_limiter(delta_u)
        case ('DEB')
           call compute_update(delta_u, x, y, u_eq, dudt)
           w1 = delta_u + dt * dudt
           call apply_limiter(w1)
           delta_u = w1
        end select
        t = t + dt
        iter = iter + 1
        write(*, *) 'time=', iter, t, dt, cmax
        if (make_movie .and. modulo(iter, interval) == 0) then
           call get_nodes_from_modes(delta_u, nodes, nx, ny, mx, my)
           snap_counter = snap_counter + 1
           write(filename, '(a, i5.5)') 'SIM', snap_counter
           call output_file(x, y, nodes, 1, filename)
        end if
     end do
  end if

  call get_nodes_from_modes(delta_u, nodes, nx, ny, mx, my)
  call compute_error(nodes, x, y, tend, u_anal)
  u = nodes
  iters = iter
  t_final = t
end subroutine evolve
