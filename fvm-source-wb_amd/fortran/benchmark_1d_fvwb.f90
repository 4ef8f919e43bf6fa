! benchmark_1d_fvwb — the 1D Euler-with-gravity benchmark run on MI355X through libfvwb.
!
! The driver of benchmark_1d.f90 (program main, :1-22): cell centres
! x = (i-0.5)*dx, the initial state selected by ninit (:37-64), the primitive
! equilibrium selected by nequilibrium (:134-155), and the time loop of
! evolve's 'EQL' branch (:200-261) with dt = 0.8*dx/cmax/3 and the snapshot
! format of output_file (:69-100: x, p - p_eq as '(7(1PE12.5,1X))').  The
! reference's loop passes u, not w1, to its second compute_update (:229-233),
! which makes that branch forward Euler; this driver passes w1 (Heun, as the
! 2D driver's evolve).
!
! usage: benchmark_1d_fvwb [stage|wb1|fvm|resident] [tend] [prefix]
!   stage     the loop above on the host; every residual and wave-speed
!             evaluation is the drop-in compute_update / compute_max_speed
!   wb1       the same loop with evolve's 'WB1' branch: the residual is the
!             drop-in compute_update_sr (hydrostatic reconstruction), whose
!             second call takes w1 in the reference too
!   fvm       the same loop with the plain residual of the 'FVM' branch, the
!             drop-in compute_update_fvm (second call on w1, as in stage)
!   resident  the configured device scheme instead: MUSCL-minmod, Rusanov,
!             SSP-RK2, solid walls, dt = 0.5*dx/cmax, the last step clipped to
!             tend, the whole run in one launch (fvwb1d_advance)
!   prefix    when given: <prefix>_init.dat, <prefix>_final.dat and <prefix>.bin
!             (int32 nx, nvar, steps; real(8) t; u0, w_eq, u)
! With make_movie set, the host loops also write simul/SIMnnn.dat every 200
! steps (:252-256; the folder must exist), x and p - p_eq per cell as
! output_file (:69-100) writes them.
program benchmark_1d_fvwb
  use iso_c_binding
  use parameters
  use fvwb_iface
  use fvwb_dropin_1d
  implicit none
  real(kind=8), allocatable :: x(:), u(:,:), u0(:,:), w_eq(:,:)
  character(len=512) :: arg, mode, prefix
  integer :: nargs, iter, ios, i
  real(kind=8) :: t, dx

  mode = 'stage'
  prefix = ''
  nargs = command_argument_count()
  if (nargs >= 1) call get_command_argument(1, mode)
  if (nargs >= 2) then
     call get_command_argument(2, arg)
     read(arg, *, iostat=ios) tend
     if (ios /= 0) error stop 'benchmark_1d_fvwb: tend must be a number'
  end if
  if (nargs >= 3) call get_command_argument(3, prefix)

  allocate(x(nx), u(nvar, nx), u0(nvar, nx), w_eq(nvar, nx))
  dx = boxlen / dble(nx)
  do i = 1, nx
     x(i) = (dble(i) - 0.5d0) * dx
  end do
  call initial_state(x, u)
  call equilibrium_primitive(x, w_eq)
  u0 = u
  if (len_trim(prefix) > 0) call write_snapshot(x, u, w_eq, trim(prefix)//'_init.dat')

  select case (trim(mode))
  case ('stage')
     call evolve_stages(u, w_eq, t, iter, 'EQL')
  case ('wb1')
     call evolve_stages(u, w_eq, t, iter, 'WB1')
  case ('fvm')
     call evolve_stages(u, w_eq, t, iter, 'FVM')
  case ('resident')
     call evolve_resident(u, w_eq, t, iter)
  case default
     error stop 'benchmark_1d_fvwb: mode must be stage, wb1, fvm or resident'
  end select
  write(*, '(a,a,a,i0,a,i0,a,1pe24.16)') 'benchmark_1d_fvwb mode=', trim(mode), ' nx=', nx, ' steps=', iter, &
       ' t=', t
  if (len_trim(prefix) > 0) then
     call write_snapshot(x, u, w_eq, trim(prefix)//'_final.dat')
     call write_state(trim(prefix)//'.bin', u0, w_eq, u, iter, t)
  end if
  call fvwb1d_dropin_close()

contains

  subroutine primitive_equilibrium_at(xv, w)
    real(kind=8), intent(in) :: xv
    real(kind=8), intent(out) :: w(nvar)
    select case (nequilibrium)
    case (3)
       w(1) = (1 - ((gamma - 1) / gamma) * 1 * xv)**(1 / (gamma - 1))
       w(2) = 0
       w(3) = (1 - ((gamma - 1) / gamma) * 1 * xv)**(gamma / (gamma - 1))
    case default
       w(1) = exp(-xv)
       w(2) = 0
       w(3) = exp(-xv)
    end select
  end subroutine primitive_equilibrium_at

  subroutine equilibrium_primitive(x, w)
    real(kind=8), intent(in) :: x(nx)
    real(kind=8), intent(out) :: w(nvar, nx)
    integer :: k
    do k = 1, nx
       call primitive_equilibrium_at(x(k), w(:, k))
    end do
  end subroutine equilibrium_primitive

  subroutine initial_state(x, u)
    real(kind=8), intent(in) :: x(nx)
    real(kind=8), intent(out) :: u(nvar, nx)
    real(kind=8) :: w(nvar, nx)
    integer :: k
    select case (ninit)
    case (1)
       w(1, :) = exp(-x)
       w(2, :) = 0
       w(3, :) = exp(-x)
    case (2)
       w(1, :) = exp(-x)
       w(2, :) = 0
       w(3, :) = exp(-x) + eta * exp(-100 * (x - boxlen / 2.)**2)
    case (3)
       do k = 1, nx
          w(1, k) = (1 - ((gamma - 1) / gamma) * 1 * x(k))**(1 / (gamma - 1))
          w(2, k) = 0
          w(3, k) = (1 - ((gamma - 1) / gamma) * 1 * x(k))**(gamma / (gamma - 1))
       end do
    case default
       error stop 'benchmark_1d_fvwb: unknown ninit'
    end select
    u(1, :) = w(1, :)
    u(2, :) = w(1, :) * w(2, :)
    u(3, :) = w(3, :) / (gamma - 1.0) + 0.5 * w(1, :) * w(2, :)**2
  end subroutine initial_state

  subroutine evolve_stages(u, w_eq, t, iter, which)
    real(kind=8), intent(inout) :: u(nvar, nx)
    real(kind=8), intent(in) :: w_eq(nvar, nx)
    real(kind=8), intent(out) :: t
    integer, intent(out) :: iter
    character(len=3), intent(in) :: which
    real(kind=8), allocatable :: dudt(:,:), w1(:,:)
    real(kind=8) :: dt, cmax
    integer :: snap_counter
    character(len=16) :: filename
    allocate(dudt(nvar, nx), w1(nvar, nx))
    t = 0
    iter = 0
    snap_counter = 0
    do while (t < tend)
       call compute_max_speed(u, cmax)
       dt = 0.8 * dx / cmax / (2.0 * dble(1) + 1.0)
       if (which == 'WB1') then
          call compute_update_sr(u, w_eq, dudt)
          w1 = u + dt * dudt
          call compute_update_sr(w1, w_eq, dudt)
       else if (which == 'FVM') then
          call compute_update_fvm(u, w_eq, dudt)
          w1 = u + dt * dudt
          call compute_update_fvm(w1, w_eq, dudt)
       else
          call compute_update(u, w_eq, dudt)
          w1 = u + dt * dudt
          call compute_update(w1, w_eq, dudt)
       end if
       u = 0.5 * u + 0.5 * w1 + 0.5 * dt * dudt
       t = t + dt
       iter = iter + 1
       if (make_movie .and. modulo(iter, 200) == 0) then
          snap_counter = snap_counter + 1
          write(filename, '(a, i3.3)') 'SIM', snap_counter
          call write_snapshot(x, u, w_eq, 'simul/'//trim(filename)//'.dat')
       end if
    end do
  end subroutine evolve_stages

  subroutine evolve_resident(u, w_eq, t, iter)
    real(kind=8), intent(inout) :: u(nvar, nx)
    real(kind=8), intent(in) :: w_eq(nvar, nx)
    real(kind=8), intent(out) :: t
    integer, intent(out) :: iter
    type(c_ptr) :: h
    type(fvwb1d_params) :: p(1)
    real(c_double) :: dtl
    integer(c_int32_t) :: steps
    call fvwb1d_check(fvwb1d_params_reference(p(1), int(nx, c_int32_t), int(nequilibrium, c_int32_t)), &
         c_null_ptr, 'fvwb1d_params_reference')
    p(1)%gamma = gamma
    p(1)%boxlen = boxlen
    p(1)%bc = FVWB1D_BC_WALL
    p(1)%recon = FVWB_RECON_MUSCL_MINMOD
    p(1)%integrator = FVWB_RK2_HEUN
    p(1)%cfl = 0.5d0
    call fvwb1d_check(fvwb1d_create(h, 1_c_int32_t, p), c_null_ptr, 'fvwb1d_create')
    call fvwb1d_check(fvwb1d_upload(h, 0_c_int32_t, u, w_eq), h, 'fvwb1d_upload')
    call fvwb1d_check(fvwb1d_advance(h, huge(steps), tend), h, 'fvwb1d_advance')
    call fvwb1d_check(fvwb1d_download(h, 0_c_int32_t, u), h, 'fvwb1d_download')
    call fvwb1d_check(fvwb1d_get_time(h, 0_c_int32_t, t, dtl, steps), h, 'fvwb1d_get_time')
    call fvwb1d_check(fvwb1d_destroy(h), c_null_ptr, 'fvwb1d_destroy')
    iter = steps
  end subroutine evolve_resident

  subroutine write_snapshot(x, u, w_eq, path)
    real(kind=8), intent(in) :: x(nx), u(nvar, nx), w_eq(nvar, nx)
    character(len=*), intent(in) :: path
    integer :: k, unit
    real(kind=8) :: p
    open(newunit=unit, file=path, form='formatted', status='replace')
    do k = 1, nx
       p = (gamma - 1.0) * (u(3, k) - 0.5 * u(1, k) * (u(2, k) / u(1, k))**2)
       write(unit, '(7(1PE12.5,1X))') x(k), p - w_eq(3, k)
    end do
    close(unit)
  end subroutine write_snapshot

  subroutine write_state(path, u0, w_eq, u, iter, t)
    character(len=*), intent(in) :: path
    real(kind=8), intent(in) :: u0(nvar, nx), w_eq(nvar, nx), u(nvar, nx), t
    integer, intent(in) :: iter
    integer :: unit
    open(newunit=unit, file=path, access='stream', form='unformatted', status='replace')
    write(unit) int(nx, c_int32_t), int(nvar, c_int32_t), int(iter, c_int32_t)
    write(unit) t
    write(unit) u0, w_eq, u
    close(unit)
  end subroutine write_state

end program benchmark_1d_fvwb
