!=======================================================================
! The device drop-in for the reference's step routines.  Each subroutine
! keeps the name and argument meaning of the routine it replaces and runs
! it on the GPU through guacho_capi:
!   boundaryI     -> gm_boundary_u   (src/boundaries.f90:44-247)
!   calcprim      -> gm_download_primit (src/hydro_core.f90:245-320)
!   get_timestep  -> gm_cfl_dtp + the ramp and clamp of
!                    src/hydro_core.f90:677-694, evaluated here
!   tstep         -> gm_tstep        (src/hydro_solver.f90:134-229)
! Errors print the library's message and stop, as the reference stops on
! its own errors.
!=======================================================================
module gm_step
  use iso_c_binding
  use guacho_capi
  use parameters
  use globals
  implicit none
  type(c_ptr), save :: gm_ctx = c_null_ptr

contains

  subroutine gm_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= GM_OK) then
      print '(a,a,a,i0,a,a)', 'guacho_amd: ', what, ' failed (', rc, '): ', gm_error_string(gm_ctx)
      stop 1
    end if
  end subroutine gm_check

  !> create the device subdomain and upload u (after initflow)
  subroutine gm_device_init()
    type(gm_params) :: p
    type(gm_conduction_params) :: tp
    integer :: dev
    character(len=16) :: env
    integer :: envlen, envstat
    dev = 0
    call get_environment_variable('GUACHO_DEVICE', env, envlen, envstat)
    if (envstat == 0 .and. envlen > 0) read(env, *) dev
    p%nx = nx; p%ny = ny; p%nz = nz
    p%neq = neq
    p%nghost = nghost
    p%riemann_solver = riemann_solver
    p%slope_limiter = slope_limiter
    p%bc = (/ bc_left, bc_right, bc_bottom, bc_top, bc_out, bc_in /)
    p%mhd = merge(1, 0, mhd)
    p%pmhd = merge(1, 0, pmhd)
    p%enable_flux_cd = merge(1, 0, enable_flux_cd)
    p%eight_wave = merge(1, 0, eight_wave)
    p%precision = GM_PRECISION_FP64
    p%device = dev
    p%reserved = 0
    p%dx = dx; p%dy = dy; p%dz = dz
    p%cv = cv
    p%gamma = gamma
    p%eta = eta
    if ((cooling /= COOL_NONE .and. cooling /= COOL_H .and. cooling /= COOL_DMC .and. cooling /= COOL_CHI) &
        .or. dif_rad .or. eq_of_state == EOS_CHEM) then
      print '(a)', 'guacho_amd: BBC and chemistry cooling, diffuse radiation and the chemistry EOS are not built'
      stop 1
    end if
    if (radiation_pressure) then
      print '(a)', 'guacho_amd: radiation pressure (a PASSIVES source term) is not built'
      stop 1
    end if
    call gm_check(gm_create(p, gm_ctx), 'gm_create')
    call gm_check(gm_set_eos(gm_ctx, int(eq_of_state, c_int), real(Tempsc, c_double)), 'gm_set_eos')
    if (cooling == COOL_DMC .or. cooling == COOL_CHI) call load_cooling()
    if (cooling == COOL_H) then
      ! coolingh takes no table; a zero table stands for the unused argument
      block
        real(c_double) :: notab(2, 41)
        notab = 0.0_c_double
        call gm_check(gm_set_cooling(gm_ctx, int(COOL_H, c_int), notab, real(Tempsc, c_double), &
                                     real(Psc, c_double), real(tsc, c_double)), 'gm_set_cooling')
      end block
    end if
    if (th_cond /= TC_OFF) then
      tp%kind = th_cond
      tp%saturation = merge(1, 0, tc_saturation)
      tp%tempsc = Tempsc; tp%rsc = rsc; tp%rhosc = rhosc; tp%mu = mu
      tp%psc = Psc; tp%tsc = tsc; tp%vsc2 = vsc2; tp%bsc = bsc
      call gm_check(gm_set_conduction(gm_ctx, tp), 'gm_set_conduction')
    end if
    if (user_source_terms) call load_user_sources()
    if (bc_user) call load_user_bc()
    call gm_check(gm_upload_u(gm_ctx, u), 'gm_upload_u')
  end subroutine gm_device_init

  !> the cooling table as read_table_dmc / read_table_chianti read it
  !> (src/cooling_dmc.f90:57-66: 10**a, 10**(-b) for DMC; raw for CHIANTI)
  subroutine load_cooling()
    real(c_double) :: tab(2, 41)
    real(kind=8) :: a, b
    integer :: i
    character(len=256) :: path
    if (cooling == COOL_DMC) then
      path = trim(workdir)//'../src/DMClib/coolingDMC.tab'
    else
      path = trim(workdir)//'../src/CHIANTIlib/coolingCHIANTI.tab'
    end if
    open(unit=10, file=trim(path), status='old')
    do i = 1, 41
      read(10, *) a, b
      if (cooling == COOL_DMC) then
        tab(1, i) = 10.e0**(a)
        tab(2, i) = 10.e0**(-b)
      else
        tab(1, i) = a
        tab(2, i) = b
      end if
    end do
    close(unit=10)
    call gm_check(gm_set_cooling(gm_ctx, int(cooling, c_int), tab, Tempsc, Psc, tsc), 'gm_set_cooling')
  end subroutine load_cooling

  !> user_source_terms: get_user_source_terms runs on the device as a
  !> plugin kernel; GUACHO_USER_SOURCES names it as <code object>:<kernel>
  subroutine load_user_sources()
    character(len=1024) :: spec
    integer :: n, st, colon
    call get_environment_variable('GUACHO_USER_SOURCES', spec, n, st)
    colon = index(spec(1:max(n, 0)), ':', back=.true.)
    if (st /= 0 .or. n == 0 .or. colon < 2 .or. colon == n) then
      print '(a)', 'guacho_amd: user_source_terms needs the device plugin: '// &
                   'set GUACHO_USER_SOURCES=<code object>:<kernel>'
      stop 1
    end if
    call gm_check(gm_set_user_sources(gm_ctx, spec(1:colon-1)//c_null_char, spec(colon+1:n)//c_null_char), &
                  'gm_set_user_sources')
  end subroutine load_user_sources

  !> bc_user: impose_user_bc runs on the device as a plugin kernel;
  !> GUACHO_USER_BC names it as <code object>:<kernel>
  subroutine load_user_bc()
    character(len=1024) :: spec
    integer :: n, st, colon
    call get_environment_variable('GUACHO_USER_BC', spec, n, st)
    colon = index(spec(1:max(n, 0)), ':', back=.true.)
    if (st /= 0 .or. n == 0 .or. colon < 2 .or. colon == n) then
      print '(a)', 'guacho_amd: bc_user needs the device plugin: set GUACHO_USER_BC=<code object>:<kernel>'
      stop 1
    end if
    call gm_check(gm_set_user_bc(gm_ctx, spec(1:colon-1)//c_null_char, spec(colon+1:n)//c_null_char), &
                  'gm_set_user_bc')
  end subroutine load_user_bc

  subroutine gm_device_finalize()
    integer(c_int) :: rc
    if (c_associated(gm_ctx)) rc = gm_destroy(gm_ctx)
    gm_ctx = c_null_ptr
  end subroutine gm_device_finalize

  !> host copy of u (ghosts after boundaryI), e.g. for dumps
  subroutine gm_download()
    call gm_check(gm_download_u(gm_ctx, u), 'gm_download_u')
  end subroutine gm_download

  subroutine boundaryI()
    call gm_check(gm_set_time(gm_ctx, time), 'gm_set_time')
    call gm_check(gm_boundary_u(gm_ctx), 'gm_boundary_u')
  end subroutine boundaryI

  !> primitives of the device state into primit (and Temp)
  subroutine calcprim()
    call gm_check(gm_download_primit(gm_ctx, primit, Temp), 'gm_download_primit')
  end subroutine calcprim

  subroutine get_timestep(current_iter, n_iter, current_time, tprint, dt, dump_flag)
    integer, intent(in) :: current_iter, n_iter
    real, intent(in) :: current_time, tprint
    real, intent(out) :: dt
    logical, intent(inout) :: dump_flag
    real(c_double) :: dtp
    call gm_check(gm_cfl_dtp(gm_ctx, dtp), 'gm_cfl_dtp')
    if (current_iter <= n_iter) then
      dtp = cfl * (2.**(-real(n_iter + 1 - current_iter))) * dtp
    else
      dtp = cfl * dtp
    end if
    dt = dtp
    if ((current_time + dt) >= tprint) then
      dt = tprint - current_time
      dump_flag = .true.
    end if
  end subroutine get_timestep

  subroutine tstep()
    call gm_check(gm_set_time(gm_ctx, time), 'gm_set_time')
    call gm_check(gm_tstep(gm_ctx, dt_CFL), 'gm_tstep')
  end subroutine tstep

end module gm_step
