!=======================================================================
! Problem plug-in of the checker's test problem: the four hooks every
! Guacho problem provides (interface as OT/user_mod.f90:43-122), all empty
! because the checker sets the state through oracle/ref/ref_kat.f90.
!=======================================================================
module user_mod
  implicit none
contains

  subroutine init_user_mod()
  end subroutine init_user_mod

  subroutine initial_conditions(u)
    use parameters, only : neq, nxmin, nxmax, nymin, nymax, nzmin, nzmax
    real, intent(out) :: u(neq, nxmin:nxmax, nymin:nymax, nzmin:nzmax)
    u = 0.
  end subroutine initial_conditions

  subroutine impose_user_bc(u, order)
    use parameters, only : neq, nxmin, nxmax, nymin, nymax, nzmin, nzmax
    real :: u(neq, nxmin:nxmax, nymin:nymax, nzmin:nzmax)
    integer, intent(in) :: order
  end subroutine impose_user_bc

  ! The checker's user source (active in builds with user_source_terms):
  ! uniform gravity along -y plus a softened point mass at the box centre,
  ! in the style of EXO/user_mod.f90:159-200 (r^3 written as r2*sqrt(r2)).
  subroutine get_user_source_terms(pp, s, i, j, k)
    use parameters, only : neq
    use globals, only : dx, dy, dz
    integer, intent(in) :: i, j, k
    real, intent(in) :: pp(neq)
    real, intent(inout) :: s(neq)
    real, parameter :: g0 = 0.5, gm = 0.02, soft = 0.01
    real :: xc, yc, zc, r2, r3
    xc = (real(i) - 0.5) * dx - 0.5
    yc = (real(j) - 0.5) * dy - 0.5
    zc = (real(k) - 0.5) * dz - 0.5
    r2 = xc * xc + yc * yc + zc * zc + soft
    r3 = r2 * sqrt(r2)
    s(3) = s(3) - pp(1) * g0
    s(5) = s(5) - pp(1) * pp(3) * g0
    s(2) = s(2) - pp(1) * gm * xc / r3
    s(3) = s(3) - pp(1) * gm * yc / r3
    s(4) = s(4) - pp(1) * gm * zc / r3
    s(5) = s(5) - pp(1) * gm * (pp(2) * xc + pp(3) * yc + pp(4) * zc) / r3
  end subroutine get_user_source_terms

end module user_mod
