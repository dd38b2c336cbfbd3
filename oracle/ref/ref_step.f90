!=======================================================================
! C-callable entry points into the reference's own time step
! (src/hydro_solver.f90: tstep, step, viscous_copy), for the oracle.  Test
! infrastructure only: oracle/Makefile links this file into the "step"
! variants of oracle/_ref, whose parameters set enable_flux_cd = .false.
!
! Those variants compile src/flux_cd_module.f90 and src/hydro_solver.f90
! without BFIELD: in hydro_solver.f90 that macro only guards the flux-CD
! calls (get_efield at :95-97, flux_cd_update at :109-111), which are dead
! code when enable_flux_cd is false, so the compiled step and tstep are the
! reference's own for that configuration.  (flux_cd_module.f90 itself does
! not compile serially: its non-MPI branch leaves an `if` open at :120.)
!=======================================================================
#if REF_FLUX_CD
#error "ref_step.f90 is only built with REF_FLUX_CD = 0"
#endif
module ref_step
  use iso_c_binding
  use parameters
  use globals
  use hydro_solver, only : tstep, step, viscous_copy
  implicit none

contains

  ! one tstep() with globals::dt_CFL = dt (hydro_solver.f90:134-229)
  subroutine ref_tstep(dt) bind(C, name='ref_tstep')
    real(c_double), value :: dt
    dt_CFL = dt
    call tstep()
  end subroutine ref_tstep

  ! step(dt) (hydro_solver.f90:77-127) on the current u, f, g, h, primit
  subroutine ref_step(dt) bind(C, name='ref_step')
    real(c_double), value :: dt
    call step(dt)
  end subroutine ref_step

  ! viscous_copy (hydro_solver.f90:47-65): u = up + eta * (laplacian of up)
  subroutine ref_viscous_copy() bind(C, name='ref_viscous_copy')
    call viscous_copy()
  end subroutine ref_viscous_copy

  ! enable_flux_cd (always 0 here) and eta of this build
  subroutine ref_step_config(fcd, etav) bind(C, name='ref_step_config')
    integer(c_int), intent(out) :: fcd
    real(c_double), intent(out) :: etav
    fcd = merge(1, 0, enable_flux_cd)
    etav = eta
  end subroutine ref_step_config

end module ref_step
