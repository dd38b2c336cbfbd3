!=======================================================================
! BIN dumps in the format of write_BIN (src/Out_BIN_Module.f90:40-248):
! ASCII header ending in achar(255), kind 'd', binary header, then
! u(:,:,:,:) including ghosts; optionally divB(nx,ny,nz) with the dump
! formula of :214-222.  The host copy of u is refreshed from the device
! first.
!=======================================================================
module out_bin
  use parameters
  use globals
  implicit none
contains

  subroutine write_header(unit, neq_out, nghost_out)
    integer, intent(in) :: unit, neq_out, nghost_out
    character, parameter :: lf = char(10)
    character(len=128) :: cbuffer
    write(unit) "**************** Output for Guacho v1.3****************", lf
    write(cbuffer, '("Dimensions    : ", i0,1x,i0,1x,i0)') nx, ny, nz
    write(unit) trim(cbuffer), lf
    write(cbuffer, '("Spacings      : ", 3(es10.3))') dx, dy, dz
    write(unit) trim(cbuffer), lf
    write(cbuffer, '("Block Origin, cells    : ", i0,1x,i0,1x,i0)') coords(0)*nx, coords(1)*ny, coords(2)*nz
    write(unit) trim(cbuffer), lf
    write(cbuffer, '("MPI blocks (X, Y, Z)   : ", i0,1x,i0,1x,i0)') MPI_NBX, MPI_NBY, MPI_NBZ
    write(unit) trim(cbuffer), lf
    write(cbuffer, '("Number of Equations/dynamical ones  ", i0,"/",i0)') neq_out, neqdyn
    write(unit) trim(cbuffer), lf
    write(cbuffer, '("Number of Ghost Cells  ", i0)') nghost_out
    write(unit) trim(cbuffer), lf
    write(cbuffer, '("Scalings ", a )')
    write(unit) trim(cbuffer), lf
    write(cbuffer, '("r_sc: ",es10.3," v_sc: ",es10.3," rho_sc: ",es10.3)') rsc, vsc, rhosc
    write(unit) trim(cbuffer), lf
    write(cbuffer, '("Specfic heat at constant volume Cv: ",f7.2)') cv
    write(unit) trim(cbuffer), lf
    write(unit) "Double precision 8 byte floats", lf
    write(unit) "*******************************************************", lf
    write(unit) achar(255), lf
    write(unit) 'd'
    write(unit) nx, ny, nz
    write(unit) dx, dy, dz
    write(unit) coords(0)*nx, coords(1)*ny, coords(2)*nz
    write(unit) MPI_NBX, MPI_NBY, MPI_NBZ
    write(unit) neq_out, neqdyn
    write(unit) nghost_out
    write(unit) rsc, vsc, rhosc
    write(unit) cv
  end subroutine write_header

  subroutine write_BIN(itprint)
    use gm_step, only : gm_download
    integer, intent(in) :: itprint
    character(len=128) :: file1
    integer :: unitout, i, j, k
    real, allocatable :: divB(:,:,:)
    call gm_download()
    write(file1, '(a,i3.3,a)') trim(outputpath)//'BIN/points', itprint, '.bin'
    open(newunit=unitout, file=file1, status='replace', access='stream')
    call write_header(unitout, neq, nghost)
    write(unitout) u(:,:,:,:)
    close(unitout)
    print '(i3,a,a)', rank, ' wrote file:', trim(file1)
    if (dump_divb) then
      allocate(divB(nx, ny, nz))
      do k = 1, nz
        do j = 1, ny
          do i = 1, nx
            divB(i,j,k) = (u(6,i+1,j,k) - u(6,i-1,j,k)) / (2.*dx) + &
                          (u(7,i,j+1,k) - u(7,i,j-1,k)) / (2.*dy) + &
                          (u(8,i,j,k+1) - u(8,i,j,k-1)) / (2.*dz)
          end do
        end do
      end do
      write(file1, '(a,i3.3,a)') trim(outputpath)//'BIN/divB-', itprint, '.bin'
      open(newunit=unitout, file=file1, status='replace', access='stream')
      call write_header(unitout, 1, 0)
      write(unitout) divB(:,:,:)
      close(unitout)
      deallocate(divB)
    end if
  end subroutine write_BIN

end module out_bin
