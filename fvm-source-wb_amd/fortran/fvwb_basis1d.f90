! fvwb_basis1d — the 1D DG basis and Gauss rule as the reference defines them
! (legendre.f90:1-127), for the driver's set-up and reconstruction: phi_0..phi_2
! normalised with square roots of default-real literals, phi_3.. the plain
! Legendre polynomials, the argument clamped to [-1,1]; the Gauss points from
! literals up to three points, else 50 Newton steps, with the weight written
! 2(2n+1)/((1-x^2) P_n'^2).  The expressions keep the reference's literals
! because the basis is part of what a mode means.
module fvwb_basis1d
  implicit none
  private
  public :: basis1d, basis1d_prime, gauss1d

contains

  real(kind=8) function basis1d(xin, k)
    real(kind=8), intent(in) :: xin
    integer, intent(in) :: k
    real(kind=8) :: x
    x = min(max(xin, -1.0), 1.0)
    select case (k)
    case (0)
       basis1d = 1.0 * sqrt(0.5)
    case (1)
       basis1d = x * 0.5 * sqrt(6.)
    case (2)
       basis1d = 0.25 * (3.0 * x**2 - 1.0) * sqrt(10.)
    case (3)
       basis1d = 0.5 * (5.0 * x**3 - 3.0 * x)
    case (4)
       basis1d = 0.125 * (35.0 * x**4 - 30.0 * x**2 + 3.0)
    case (5)
       basis1d = 0.125 * (63.0 * x**5 - 70.0 * x**3 + 15.0 * x)
    case default
       basis1d = 1.0 / 16.0 * (231.0 * x**6 - 315.0 * x**4 + 105.0 * x**2 - 5.0)
    end select
  end function basis1d

  real(kind=8) function basis1d_prime(xin, k)
    real(kind=8), intent(in) :: xin
    integer, intent(in) :: k
    real(kind=8) :: x
    x = min(max(xin, -1.0), 1.0)
    select case (k)
    case (0)
       basis1d_prime = 0.0
    case (1)
       basis1d_prime = 1.0 * 0.5 * sqrt(6.)
    case (2)
       basis1d_prime = 6.0 * x * 0.25 * sqrt(10.)
    case (3)
       basis1d_prime = 0.5 * (15.0 * x**2 - 3.0)
    case (4)
       basis1d_prime = 0.125 * (140.0 * x**3 - 60.0 * x)
    case (5)
       basis1d_prime = 0.125 * (315.0 * x**4 - 210.0 * x**2 + 15.0)
    case default
       basis1d_prime = 1.0 / 16.0 * (1386.0 * x**5 - 1260.0 * x**3 + 210.0 * x)
    end select
  end function basis1d_prime

  subroutine gauss1d(x, w, n)
    integer, intent(in) :: n
    real(kind=8), intent(out) :: x(n), w(n)
    integer :: i, it
    real(kind=8) :: xx, dpi
    select case (n)
    case (1)
       x(1) = 0.0
       w(1) = 2.0
    case (2)
       x(1) = -1./dble(3)*sqrt(3.)
       x(2) = 1./dble(3)*sqrt(3.)
       w = 1.
    case (3)
       x(1) = -sqrt(3.)/sqrt(5.)
       x(2) = 0.0
       x(3) = sqrt(3.)/sqrt(5.)
       w(1) = 5./9.
       w(2) = 8./9.
       w(3) = 5./9.
    case default
       dpi = acos(-1.0d0)
       do i = 1, n
          xx = (1.0 - 0.125/n/n + 0.125/n/n/n) * cos(dpi * (4.0 * dble(i) - 1.0) / (4.0 * dble(n) + 2.0))
          do it = 1, 50
             xx = xx - basis1d(xx, n) / basis1d_prime(xx, n)
          end do
          x(i) = xx
          w(i) = 2.0 * (2.0 * dble(n) + 1.0) / (1.0 - x(i)**2) / basis1d_prime(x(i), n)**2
       end do
       do i = n/2 + 1, n
          x(i) = -x(n - i + 1)
          w(i) = w(n - i + 1)
       end do
    end select
  end subroutine gauss1d

end module fvwb_basis1d
