! fvwb_cpu_fv2d — CPU FP64 restatement, in Fortran, of the configured 2D
! finite-volume scheme (TEST INFRASTRUCTURE: the timed CPU baseline and a
! full-size checker; never linked by the product).
!
! The scheme is the one of oracle/fvwb_oracle_fv2d.c (orc_fv2d_residual,
! orc_fv2d_step, orc_fv2d_advance) for reflecting walls, MUSCL with the
! limiters.f90:1-21 minmod, HLL with Davis bounds and SSP-RK3: the
! deviation du = u - U_eq is reconstructed, the face flux minus the
! equilibrium flux at the face point is differenced, and the source minus the
! equilibrium source (benchmark_2d.f90:327-350, grad phi = (g, g)) is added.
! Statements round in the same order as the C restatement, so built with
! -ffp-contract=off the two agree bit for bit.
!
! Arrays are stored as the reference stores them, (nvar, nx, ny).  Unlike the
! reference's compute_update_exact (benchmark_2d.f90:465-618), which evaluates
! the equilibrium at every face on every call, the conservative equilibrium at
! the x and y faces is tabulated once before the time loop: the per-call work
! is then the residual proper.
!
! Usage: fvwb_cpu_fv2d nx ny steps gamma cfl rho0 p0 g infile outfile [threads]
!   infile : u then w_eq, each (4, nx, ny) float64 in Fortran order
!   outfile: u after `steps` SSP-RK3 steps, same layout ('-' writes nothing)
! Prints "seconds <s> steps <n> t <t> dt <dt>" (wall clock of the time loop,
! system_clock).  With OpenMP, the row loops run on `threads` threads
! (default 1).
module cpu_fv2d
  implicit none
  integer, parameter :: dp = kind(1.0d0)
  integer :: nx, ny
  real(dp) :: gamma, cfl, rho0, p0, gg, dx, dy, odx, ody
  real(dp), allocatable :: uex(:,:,:), uey(:,:,:)      ! conservative equilibrium at x and y faces
  real(dp), allocatable :: fex(:,:,:), fey(:,:,:)      ! its x flux at x faces, y flux at y faces
contains

  subroutine to_primitive(u, w)
    real(dp), intent(in) :: u(4)
    real(dp), intent(out) :: w(4)
    w(1) = u(1)
    w(2) = u(2) / w(1)
    w(3) = u(3) / w(1)
    w(4) = (gamma - 1.0_dp) * (u(4) - 0.5_dp * w(1) * (w(2) * w(2) + w(3) * w(3)))
  end subroutine to_primitive

  subroutine to_conservative(w, u)
    real(dp), intent(in) :: w(4)
    real(dp), intent(out) :: u(4)
    u(1) = w(1)
    u(2) = w(1) * w(2)
    u(3) = w(1) * w(3)
    u(4) = w(4) / (gamma - 1.0_dp) + 0.5_dp * (w(1) * (w(2) * w(2) + w(3) * w(3)))
  end subroutine to_conservative

  ! physical flux along direction dir (0: x, 1: y), benchmark_2d.f90:299-325
  subroutine phys_flux(dir, u, f)
    integer, intent(in) :: dir
    real(dp), intent(in) :: u(4)
    real(dp), intent(out) :: f(4)
    real(dp) :: w(4)
    call to_primitive(u, w)
    if (dir == 0) then
      f(1) = w(2) * u(1)
      f(2) = w(2) * u(2) + w(4)
      f(3) = w(1) * w(2) * w(3)
      f(4) = w(2) * u(4) + w(2) * w(4)
    else
      f(1) = u(1) * w(3)
      f(2) = u(2) * w(3)
      f(3) = u(3) * w(3) + w(4)
      f(4) = w(3) * u(4) + w(3) * w(4)
    end if
  end subroutine phys_flux

  ! limiters.f90:1-21: sign taken from the sign bit (sign(1,x) is +1 at +0)
  pure real(dp) function minmod3(a, b, c)
    real(dp), intent(in) :: a, b, c
    real(dp) :: s
    s = sign(1.0_dp, a)
    if (sign(1.0_dp, b) == s .and. sign(1.0_dp, c) == s) then
      minmod3 = s * min(min(abs(a), abs(b)), abs(c))
    else
      minmod3 = 0.0_dp
    end if
  end function minmod3

  ! deviation of cell (i, j), 0-based, mirrored at the walls with the normal
  ! momentum negated
  subroutine dev_at(du, i0, j0, d)
    real(dp), intent(in) :: du(4, 0:nx-1, 0:ny-1)
    integer, intent(in) :: i0, j0
    real(dp), intent(out) :: d(4)
    integer :: i, j
    real(dp) :: sx, sy
    i = i0; j = j0; sx = 1.0_dp; sy = 1.0_dp
    if (i < 0) then
      i = -1 - i; sx = -1.0_dp
    else if (i >= nx) then
      i = 2 * nx - 1 - i; sx = -1.0_dp
    end if
    if (j < 0) then
      j = -1 - j; sy = -1.0_dp
    else if (j >= ny) then
      j = 2 * ny - 1 - j; sy = -1.0_dp
    end if
    d(1) = du(1, i, j); d(2) = sx * du(2, i, j); d(3) = sy * du(3, i, j); d(4) = du(4, i, j)
  end subroutine dev_at

  ! HLL flux minus the equilibrium flux at a face with conservative
  ! equilibrium ue (flux fe) and face deviations dl, dr
  subroutine face_flux(dir, ue, fe, dl, dr, out)
    integer, intent(in) :: dir
    real(dp), intent(in) :: ue(4), fe(4), dl(4), dr(4)
    real(dp), intent(out) :: out(4)
    real(dp) :: ul(4), ur(4), fl(4), fr(4), wl(4), wr(4)
    real(dp) :: cl, cr, vl, vr, sl, sr, ap, am, inv, wgt, q
    integer :: k
    do k = 1, 4
      ul(k) = ue(k) + dl(k)
      ur(k) = ue(k) + dr(k)
    end do
    call phys_flux(dir, ul, fl)
    call phys_flux(dir, ur, fr)
    call to_primitive(ul, wl)
    call to_primitive(ur, wr)
    cl = sqrt(gamma * max(wl(4), 1e-10_dp) / max(wl(1), 1e-10_dp))
    cr = sqrt(gamma * max(wr(4), 1e-10_dp) / max(wr(1), 1e-10_dp))
    vl = wl(2 + dir); vr = wr(2 + dir)
    sl = min(vl - cl, vr - cr); sr = max(vl + cl, vr + cr)
    ap = max(0.0_dp, sr); am = max(0.0_dp, -sl)
    inv = 1.0_dp / (ap + am)
    wgt = am * inv; q = ap * am * inv
    do k = 1, 4
      out(k) = (fl(k) + wgt * (fr(k) - fl(k)) - q * (ur(k) - ul(k))) - fe(k)
    end do
  end subroutine face_flux

  ! MUSCL face deviations between cell a = (ia, ja) and the next cell along dir
  subroutine face_devs(du, dir, ia, ja, dl, dr)
    real(dp), intent(in) :: du(4, 0:nx-1, 0:ny-1)
    integer, intent(in) :: dir, ia, ja
    real(dp), intent(out) :: dl(4), dr(4)
    real(dp) :: a(4), b(4), am(4), bp(4), sa, sb
    integer :: di, dj, k
    di = 1 - dir; dj = dir
    call dev_at(du, ia, ja, a)
    call dev_at(du, ia + di, ja + dj, b)
    call dev_at(du, ia - di, ja - dj, am)
    call dev_at(du, ia + 2 * di, ja + 2 * dj, bp)
    do k = 1, 4
      sa = minmod3(a(k) - am(k), b(k) - a(k), b(k) - a(k))
      sb = minmod3(b(k) - a(k), bp(k) - b(k), bp(k) - b(k))
      dl(k) = a(k) + 0.5_dp * sa
      dr(k) = b(k) - 0.5_dp * sb
    end do
  end subroutine face_devs

  subroutine eq_point(x, y, w)
    real(dp), intent(in) :: x, y
    real(dp), intent(out) :: w(4)
    real(dp) :: k
    k = rho0 * gg / p0
    w(1) = rho0 * exp(-(k * (x + y)))
    w(2) = 0.0_dp
    w(3) = 0.0_dp
    w(4) = p0 * exp(-(k * (x + y)))
  end subroutine eq_point

  subroutine tabulate_faces()
    integer :: i, j
    real(dp) :: we(4)
    allocate(uex(4, 0:nx, 0:ny-1), uey(4, 0:nx-1, 0:ny), fex(4, 0:nx, 0:ny-1), fey(4, 0:nx-1, 0:ny))
    do j = 0, ny - 1
      do i = 0, nx
        call eq_point(dble(i) * dx, (dble(j) + 0.5_dp) * dy, we)
        call to_conservative(we, uex(:, i, j))
        call phys_flux(0, uex(:, i, j), fex(:, i, j))
      end do
    end do
    do j = 0, ny
      do i = 0, nx - 1
        call eq_point((dble(i) + 0.5_dp) * dx, dble(j) * dy, we)
        call to_conservative(we, uey(:, i, j))
        call phys_flux(1, uey(:, i, j), fey(:, i, j))
      end do
    end do
  end subroutine tabulate_faces

  ! residual of the well-balanced deviation form; du, Fx, Gy are work arrays
  subroutine residual(u, weq, ueq, du, Fx, Gy, dudt)
    real(dp), intent(in) :: u(4, 0:nx-1, 0:ny-1), weq(4, 0:nx-1, 0:ny-1), ueq(4, 0:nx-1, 0:ny-1)
    real(dp), intent(inout) :: du(4, 0:nx-1, 0:ny-1), Fx(4, 0:nx, 0:ny-1), Gy(4, 0:nx-1, 0:ny)
    real(dp), intent(out) :: dudt(4, 0:nx-1, 0:ny-1)
    real(dp) :: dl(4), dr(4), w(4), s(4), seq(4)
    integer :: i, j, k
    !$omp parallel private(i, j, k, dl, dr, w, s, seq)
    !$omp do schedule(static)
    do j = 0, ny - 1
      do i = 0, nx - 1
        do k = 1, 4
          du(k, i, j) = u(k, i, j) - ueq(k, i, j)
        end do
      end do
    end do
    !$omp end do
    !$omp do schedule(static)
    do j = 0, ny - 1
      do i = 0, nx
        call face_devs(du, 0, i - 1, j, dl, dr)
        call face_flux(0, uex(:, i, j), fex(:, i, j), dl, dr, Fx(:, i, j))
      end do
    end do
    !$omp end do
    !$omp do schedule(static)
    do j = 0, ny
      do i = 0, nx - 1
        call face_devs(du, 1, i, j - 1, dl, dr)
        call face_flux(1, uey(:, i, j), fey(:, i, j), dl, dr, Gy(:, i, j))
      end do
    end do
    !$omp end do
    !$omp do schedule(static)
    do j = 0, ny - 1
      do i = 0, nx - 1
        call to_primitive(u(:, i, j), w)
        s(1) = 0.0_dp; s(2) = -w(1) * gg; s(3) = -w(1) * gg; s(4) = -w(1) * (w(2) * gg + w(3) * gg)
        seq(1) = 0.0_dp; seq(2) = -weq(1, i, j) * gg; seq(3) = -weq(1, i, j) * gg
        seq(4) = -weq(1, i, j) * (weq(2, i, j) * gg + weq(3, i, j) * gg)
        do k = 1, 4
          dudt(k, i, j) = (-(Fx(k, i + 1, j) - Fx(k, i, j)) * odx - (Gy(k, i, j + 1) - Gy(k, i, j)) * ody) &
                          + (s(k) - seq(k))
        end do
      end do
    end do
    !$omp end do
    !$omp end parallel
  end subroutine residual

  ! benchmark_2d.f90:264-279 compute_max_speed
  real(dp) function max_speed(u)
    real(dp), intent(in) :: u(4, 0:nx-1, 0:ny-1)
    real(dp) :: w(4), cs, c
    integer :: i, j
    c = 0.0_dp
    !$omp parallel do private(i, w, cs) reduction(max:c) schedule(static)
    do j = 0, ny - 1
      do i = 0, nx - 1
        call to_primitive(u(:, i, j), w)
        cs = sqrt(gamma * max(w(4), 1e-10_dp) / max(w(1), 1e-10_dp))
        c = max(c, sqrt(w(2) * w(2) + w(3) * w(3)) + cs)
      end do
    end do
    !$omp end parallel do
    max_speed = c
  end function max_speed
  ! out = a*x + b*y + c*z, except a = 0 which gives y + c*z (the first stage,
  ! u + dt*L, rounds as written)
  subroutine combine(a, x, b, y, c, z, out)
    real(dp), intent(in) :: a, b, c
    real(dp), intent(in) :: x(4 * nx * ny), y(4 * nx * ny), z(4 * nx * ny)
    real(dp), intent(inout) :: out(4 * nx * ny)
    integer :: q
    if (a == 0.0_dp) then
      !$omp parallel do schedule(static)
      do q = 1, 4 * nx * ny
        out(q) = y(q) + c * z(q)
      end do
      !$omp end parallel do
    else
      !$omp parallel do schedule(static)
      do q = 1, 4 * nx * ny
        out(q) = a * x(q) + b * y(q) + c * z(q)
      end do
      !$omp end parallel do
    end if
  end subroutine combine
end module cpu_fv2d

program fvwb_cpu_fv2d
  use cpu_fv2d
  !$ use omp_lib
  implicit none
  character(len=512) :: arg, infile, outfile
  integer :: steps, n, i, j, threads, unit_in, unit_out
  integer(8) :: c0, c1, rate
  real(dp), allocatable :: u(:,:,:), weq(:,:,:), ueq(:,:,:), du(:,:,:), Fx(:,:,:), Gy(:,:,:)
  real(dp), allocatable :: L(:,:,:), u1(:,:,:), u2(:,:,:), tmp(:,:,:)
  real(dp) :: t, dt, h

  if (command_argument_count() < 10) then
    write(*, '(a)') 'usage: fvwb_cpu_fv2d nx ny steps gamma cfl rho0 p0 g infile outfile [threads]'
    stop 2
  end if
  call get_command_argument(1, arg); read(arg, *) nx
  call get_command_argument(2, arg); read(arg, *) ny
  call get_command_argument(3, arg); read(arg, *) steps
  call get_command_argument(4, arg); read(arg, *) gamma
  call get_command_argument(5, arg); read(arg, *) cfl
  call get_command_argument(6, arg); read(arg, *) rho0
  call get_command_argument(7, arg); read(arg, *) p0
  call get_command_argument(8, arg); read(arg, *) gg
  call get_command_argument(9, infile)
  call get_command_argument(10, outfile)
  threads = 1
  if (command_argument_count() >= 11) then
    call get_command_argument(11, arg); read(arg, *) threads
  end if
  !$ call omp_set_num_threads(threads)
  if (nx < 4 .or. ny < 4 .or. steps < 0) then
    write(*, '(a)') 'fvwb_cpu_fv2d: nx and ny must be at least 4 and steps non-negative'
    stop 2
  end if

  allocate(u(4, 0:nx-1, 0:ny-1), weq(4, 0:nx-1, 0:ny-1), ueq(4, 0:nx-1, 0:ny-1), du(4, 0:nx-1, 0:ny-1))
  allocate(Fx(4, 0:nx, 0:ny-1), Gy(4, 0:nx-1, 0:ny), L(4, 0:nx-1, 0:ny-1), u1(4, 0:nx-1, 0:ny-1), &
           u2(4, 0:nx-1, 0:ny-1))
  open(newunit=unit_in, file=trim(infile), access='stream', form='unformatted', status='old', action='read')
  read(unit_in) u
  read(unit_in) weq
  close(unit_in)

  dx = 1.0_dp / dble(nx)
  dy = 1.0_dp / dble(ny)
  odx = 1.0_dp / dx
  ody = 1.0_dp / dy
  h = min(dx, dy)
  do j = 0, ny - 1
    do i = 0, nx - 1
      call to_conservative(weq(:, i, j), ueq(:, i, j))
    end do
  end do
  call tabulate_faces()

  t = 0.0_dp
  dt = 0.0_dp
  call system_clock(c0, rate)
  do n = 1, steps
    dt = 0.5_dp * h / max_speed(u) * cfl
    ! Shu-Osher SSP-RK3 with exact 3/4, 1/4, 1/3, 2/3
    call residual(u, weq, ueq, du, Fx, Gy, L)
    call combine(0.0_dp, u, 1.0_dp, u, dt, L, u1)
    call residual(u1, weq, ueq, du, Fx, Gy, L)
    call combine(0.75_dp, u, 0.25_dp, u1, 0.25_dp * dt, L, u2)
    call residual(u2, weq, ueq, du, Fx, Gy, L)
    call combine(1.0_dp / 3.0_dp, u, 2.0_dp / 3.0_dp, u2, (2.0_dp / 3.0_dp) * dt, L, u1)
    call move_alloc(u, tmp)
    call move_alloc(u1, u)
    call move_alloc(tmp, u1)
    t = t + dt
  end do
  call system_clock(c1)
  write(*, '(a,es24.16,a,i0,a,es24.16,a,es24.16)') 'seconds ', dble(c1 - c0) / dble(rate), ' steps ', steps, &
       ' t ', t, ' dt ', dt
  if (trim(outfile) /= '-') then
    open(newunit=unit_out, file=trim(outfile), access='stream', form='unformatted', status='replace', &
         action='write')
    write(unit_out) u
    close(unit_out)
  end if
end program fvwb_cpu_fv2d
