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

  subroutine get_user_source_terms(pp, s, i, j, k)
    use parameters, only : neq
    integer, intent(in) :: i, j, k
    real, intent(in) :: pp(neq)
    real, intent(inout) :: s(neq)
  end subroutine get_user_source_terms

end module user_mod
