!=======================================================================
! Host program: the time loop of the reference driver (src/main.f90:70-137)
! with the step, the ghost fill and the CFL reduction on the GPU.  Problem
! set-up comes from the user's parameters and user_mod modules, unchanged.
!=======================================================================
program guacho_amd
  use parameters
  use globals
  use init
  use gm_step
  use out_bin
  implicit none
  integer :: itprint
  real :: tprint
  logical :: dump_out = .false.

  call initmain(tprint, itprint)
  call initflow(itprint)
  call gm_device_init()
  call boundaryI()

  if (.not. iwarm) then
    if (out_bin) call write_BIN(itprint)
    itprint = itprint + 1
  end if

  do while (time <= tmax)
    call get_timestep(currentIteration, 10, time, tprint, dt_CFL, dump_out)
    print '(a,i0,a,es12.3,a,es12.3,a,es12.3,a)', 'Iteration ', currentIteration, &
          ' | time:', time*tsc, ' | dt:', dt_CFL*tsc, ' | tprint:', tprint*tsc
    call tstep()
    time = time + dt_CFL
    if (dump_out) then
      if (out_bin) call write_BIN(itprint)
      print '(a,i4)', '****************** wrote output *************** :', itprint
      tprint = tprint + dtprint
      itprint = itprint + 1
      dump_out = .false.
    end if
    currentIteration = currentIteration + 1
  end do

  print '(a)', '--- My work here is done, have a nice day ---'
  call gm_device_finalize()
end program guacho_amd
