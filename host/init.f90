!=======================================================================
! Start-up of the host: grid spacing, clock and allocations
! (src/init.f90:119-151 of the reference) and the initial state, either
! from the user's initial_conditions hook or a warm start from a BIN dump
! (the reader semantics of src/init.f90:436-477).
!=======================================================================
module init
  use parameters
  use globals
  implicit none
contains

  subroutine initmain(tprint, itprint)
    use user_mod, only : init_user_mod
    real, intent(out) :: tprint
    integer, intent(out) :: itprint
    dx = xmax / nxtot
    dy = ymax / nytot
    dz = zmax / nztot
    currentIteration = 1
    if (.not. iwarm) then
      print '(a)', 'Starting cold'
      itprint = 0
      time = 0.
      tprint = dtprint
    else
      itprint = itprint0
      time = real(itprint) * dtprint
      print '(a,i0)', 'Warm start, from output ', itprint
      tprint = time + dtprint
    end if
    allocate(u(neq, nxmin:nxmax, nymin:nymax, nzmin:nzmax))
    allocate(primit(neq, nxmin:nxmax, nymin:nymax, nzmin:nzmax))
    allocate(Temp(nxmin:nxmax, nymin:nymax, nzmin:nzmax))
    u = 0.
    primit = 0.
    Temp = 0.
    call execute_command_line('mkdir -p '//trim(outputpath)//'BIN')
    call init_user_mod()
    print '(a,i0,a,i0,a,i0,a)', 'grid ', nx, ' x ', ny, ' x ', nz, ' cells on one GPU subdomain'
  end subroutine initmain

  subroutine initflow(itprint)
    use user_mod, only : initial_conditions
    integer, intent(inout) :: itprint
    integer :: unitin
    character(len=128) :: file1
    character :: byte_read
    integer :: nxp, nyp, nzp, x0p, y0p, z0p, mpi_xp, mpi_yp, mpi_zp, neqp, neqdynp, nghostp
    real :: dxp, dyp, dzp, scal(3), cvp
    if (.not. iwarm) then
      call initial_conditions(u)
    else
      write(file1, '(a,i3.3,a)') trim(outputpath)//'BIN/points', itprint, '.bin'
      open(newunit=unitin, file=file1, status='old', access='stream')
      byte_read = ' '
      do while (byte_read /= achar(255))
        read(unitin) byte_read
      end do
      read(unitin) byte_read
      read(unitin) byte_read
      read(unitin) nxp, nyp, nzp
      read(unitin) dxp, dyp, dzp
      read(unitin) x0p, y0p, z0p
      read(unitin) mpi_xp, mpi_yp, mpi_zp
      read(unitin) neqp, neqdynp
      read(unitin) nghostp
      read(unitin) scal(1:3)
      read(unitin) cvp
      if (nxp /= nx .or. nyp /= ny .or. nzp /= nz .or. neqp /= neq .or. nghostp /= nghost) then
        print '(a,a)', 'warm start dump does not match the grid: ', trim(file1)
        stop 1
      end if
      read(unitin) u(:,:,:,:)
      close(unitin)
      print '(a,a)', 'read: ', trim(file1)
      itprint = itprint + 1
    end if
  end subroutine initflow

end module init
