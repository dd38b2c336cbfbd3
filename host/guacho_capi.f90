!=======================================================================
! ISO_C_BINDING interface to the C ABI of include/guacho_amd.h
! (guacho_amd/libguacho_amd.so).  One interface per C entry point, same
! names and argument meaning; arrays are passed as the host's
! u(neq, nxmin:nxmax, nymin:nymax, nzmin:nzmax) in place.
!=======================================================================
module guacho_capi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: GM_OK = 0, GM_ERR_CONFIG = 1, GM_ERR_HIP = 2, &
                               GM_ERR_NONFINITE = 3, GM_ERR_OOM = 4, GM_ERR_ARG = 5
  integer(c_int32_t), parameter :: GM_PRECISION_FP64 = 0, GM_PRECISION_MIXED = 1

  type, bind(C) :: gm_params
    integer(c_int32_t) :: nx, ny, nz
    integer(c_int32_t) :: neq
    integer(c_int32_t) :: nghost
    integer(c_int32_t) :: riemann_solver
    integer(c_int32_t) :: slope_limiter
    integer(c_int32_t) :: bc(6)
    integer(c_int32_t) :: mhd
    integer(c_int32_t) :: pmhd
    integer(c_int32_t) :: enable_flux_cd
    integer(c_int32_t) :: eight_wave
    integer(c_int32_t) :: precision
    integer(c_int32_t) :: device
    integer(c_int32_t) :: reserved
    real(c_double) :: dx, dy, dz
    real(c_double) :: cv
    real(c_double) :: gamma
    real(c_double) :: eta
  end type gm_params

  ! thermal conduction settings (gm_conduction_params of include/guacho_amd.h)
  type, bind(C) :: gm_conduction_params
    integer(c_int32_t) :: kind, saturation
    real(c_double) :: tempsc, rsc, rhosc, mu, psc, tsc, vsc2, bsc
  end type gm_conduction_params

  interface
    integer(c_int) function gm_create(params, ctx) bind(C, name='gm_create')
      import :: gm_params, c_ptr, c_int
      type(gm_params), intent(in) :: params
      type(c_ptr), intent(out) :: ctx
    end function gm_create

    integer(c_int) function gm_destroy(ctx) bind(C, name='gm_destroy')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function gm_destroy

    type(c_ptr) function gm_last_error(ctx) bind(C, name='gm_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function gm_last_error

    integer(c_int) function gm_upload_u(ctx, u) bind(C, name='gm_upload_u')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: u(*)
    end function gm_upload_u

    integer(c_int) function gm_download_u(ctx, u) bind(C, name='gm_download_u')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: u(*)
    end function gm_download_u

    integer(c_int) function gm_download_primit(ctx, primit, temp) bind(C, name='gm_download_primit')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: primit(*)
      real(c_double), intent(out) :: temp(*)
    end function gm_download_primit

    integer(c_int) function gm_boundary_u(ctx) bind(C, name='gm_boundary_u')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function gm_boundary_u

    integer(c_int) function gm_cfl_dtp(ctx, dtp) bind(C, name='gm_cfl_dtp')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dtp
    end function gm_cfl_dtp

    integer(c_int) function gm_tstep(ctx, dt) bind(C, name='gm_tstep')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
    end function gm_tstep

    integer(c_int) function gm_sync(ctx) bind(C, name='gm_sync')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function gm_sync

    integer(c_int) function gm_download_divb(ctx, divb) bind(C, name='gm_download_divb')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: divb(*)
    end function gm_download_divb

    integer(c_int) function gm_set_time(ctx, time) bind(C, name='gm_set_time')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: time
    end function gm_set_time

    ! code_object and kernel: NUL-terminated strings
    integer(c_int) function gm_set_user_sources(ctx, code_object, kernel) bind(C, name='gm_set_user_sources')
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: code_object(*), kernel(*)
    end function gm_set_user_sources

    integer(c_int) function gm_set_user_bc(ctx, code_object, kernel) bind(C, name='gm_set_user_bc')
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: code_object(*), kernel(*)
    end function gm_set_user_bc

    integer(c_int) function gm_set_cooling(ctx, kind, cooltab, tempsc, psc, tsc) bind(C, name='gm_set_cooling')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: kind
      real(c_double), intent(in) :: cooltab(*)
      real(c_double), value :: tempsc, psc, tsc
    end function gm_set_cooling

    integer(c_int) function gm_set_conduction(ctx, params) bind(C, name='gm_set_conduction')
      import :: c_ptr, c_int, gm_conduction_params
      type(c_ptr), value :: ctx
      type(gm_conduction_params), intent(in) :: params
    end function gm_set_conduction
    integer(c_int) function gm_set_eos(ctx, eos, tempsc) bind(C, name='gm_set_eos')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: eos
      real(c_double), value :: tempsc
    end function gm_set_eos
  end interface

contains

  !> Message of the last error as a Fortran string
  function gm_error_string(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n
    p = gm_last_error(ctx)
    if (.not. c_associated(p)) then
      msg = ''
      return
    end if
    call c_f_pointer(p, chars, [4096])
    n = 0
    do while (n < 4096)
      if (chars(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(chars(1:n), msg)
  end function gm_error_string

end module guacho_capi
