!=======================================================================
! C-callable entry points into the reference's own routines, for the
! oracle.  Test infrastructure only: oracle/Makefile links this
! file with the reference's unmodified sources (compiled where they lie
! under the reference tree) into oracle/_ref/<variant>/libguacho_ref.so,
! which tests/ load to validate the C restatement (oracle/guacho_oracle.c)
! and to write the fixtures under tests/golden/.
!
! Arrays cross the boundary in the reference's layout
! u(neq, nxmin:nxmax, nymin:nymax, nzmin:nzmax), column-major.
!=======================================================================
module ref_kat
  use iso_c_binding
  use parameters
  use globals
  use hydro_core, only : u2prim, prim2u, prim2f, cfast, cfastX, csound, &
                         limiter, calcprim, get_timestep, swapy, swapz
  use hlld, only : prim2fhlld, hlldfluxes
  use hlle, only : prim2fhlle, hllefluxes
  use hll,  only : prim2fhll, hllfluxes
  use hllc, only : prim2fhllc, hllcfluxes
  use boundaries, only : boundaryI, boundaryII
  use sources, only : source
  use cooling_dmc, only : init_cooling_dmc, coolingdmc, cooltab_dmc
  use cooling_chi, only : init_cooling_chianti, coolingchi, cooltab_chianti
#ifdef PASSIVES
  use cooling_H, only : coolingh
#endif
  use thermal_cond, only : thermal_conduction, get_dt_cond
  implicit none

contains

  subroutine ref_init(dxv, dyv, dzv) bind(C, name='ref_init')
    real(c_double), value :: dxv, dyv, dzv
    if (.not. allocated(u)) then
      allocate(     u(neq, nxmin:nxmax, nymin:nymax, nzmin:nzmax))
      allocate(    up(neq, nxmin:nxmax, nymin:nymax, nzmin:nzmax))
      allocate(primit(neq, nxmin:nxmax, nymin:nymax, nzmin:nzmax))
      allocate(     f(neq, nxmin:nxmax, nymin:nymax, nzmin:nzmax))
      allocate(     g(neq, nxmin:nxmax, nymin:nymax, nzmin:nzmax))
      allocate(     h(neq, nxmin:nxmax, nymin:nymax, nzmin:nzmax))
      allocate(  Temp(nxmin:nxmax, nymin:nymax, nzmin:nzmax))
    end if
    u = 0.; up = 0.; primit = 0.; f = 0.; g = 0.; h = 0.; Temp = 0.
    dx = dxv; dy = dyv; dz = dzv
    coords = 0
  end subroutine ref_init

  ! integer configuration: nx, ny, nz, neq, solver, limiter, bc(6), mhd, pmhd
  subroutine ref_config(cfg, rcfg) bind(C, name='ref_config')
    integer(c_int), intent(out) :: cfg(14)
    real(c_double), intent(out) :: rcfg(3)
    cfg = (/ nx, ny, nz, neq, riemann_solver, slope_limiter, bc_left, bc_right, &
             bc_bottom, bc_top, bc_out, bc_in, merge(1, 0, mhd), merge(1, 0, pmhd) /)
    rcfg = (/ cv, gamma, cfl /)
  end subroutine ref_config

  subroutine ref_u2prim(uu, prim, t) bind(C, name='ref_u2prim')
    real(c_double), intent(in)  :: uu(neq)
    real(c_double), intent(out) :: prim(neq), t
    call u2prim(uu, prim, t)
  end subroutine ref_u2prim

  subroutine ref_prim2u(prim, uu) bind(C, name='ref_prim2u')
    real(c_double), intent(in)  :: prim(neq)
    real(c_double), intent(out) :: uu(neq)
    call prim2u(prim, uu)
  end subroutine ref_prim2u

  subroutine ref_prim2f(prim, ff) bind(C, name='ref_prim2f')
    real(c_double), intent(in)  :: prim(neq)
    real(c_double), intent(out) :: ff(neq)
    call prim2f(prim, ff)
  end subroutine ref_prim2f

  subroutine ref_cfast(p, d, bx, by, bz, cf) bind(C, name='ref_cfast')
    real(c_double), value :: p, d, bx, by, bz
    real(c_double), intent(out) :: cf(3)
    call cfast(p, d, bx, by, bz, cf(1), cf(2), cf(3))
  end subroutine ref_cfast

  subroutine ref_cfastx(prim, c) bind(C, name='ref_cfastx')
    real(c_double), intent(in)  :: prim(neq)
    real(c_double), intent(out) :: c
    call cfastX(prim, c)
  end subroutine ref_cfastx

  subroutine ref_csound(p, d, c) bind(C, name='ref_csound')
    real(c_double), value :: p, d
    real(c_double), intent(out) :: c
    call csound(p, d, c)
  end subroutine ref_csound

  subroutine ref_limiter(pll, pl, pr, prr) bind(C, name='ref_limiter')
    real(c_double), intent(in)    :: pll(neq), prr(neq)
    real(c_double), intent(inout) :: pl(neq), pr(neq)
    call limiter(pll, pl, pr, prr, neq)
  end subroutine ref_limiter

  subroutine ref_swap(axis, v) bind(C, name='ref_swap')
    integer(c_int), value :: axis
    real(c_double), intent(inout) :: v(neq)
    if (axis == 1) call swapy(v, neq)
    if (axis == 2) call swapz(v, neq)
  end subroutine ref_swap

  ! one interface solve with the configured solver
  subroutine ref_riemann(pl, pr, ff) bind(C, name='ref_riemann')
    real(c_double), intent(in)    :: pl(neq), pr(neq)
    real(c_double), intent(inout) :: ff(neq)
    if (riemann_solver == SOLVER_HLLD) call prim2fhlld(pl, pr, ff)
    if (riemann_solver == SOLVER_HLLE) call prim2fhlle(pl, pr, ff)
    if (riemann_solver == SOLVER_HLL ) call prim2fhll (pl, pr, ff)
    if (riemann_solver == SOLVER_HLLC) call prim2fhllc(pl, pr, ff)
  end subroutine ref_riemann

  ! which: 0 u, 1 up, 2 primit
  subroutine ref_set_array(which, a) bind(C, name='ref_set_array')
    integer(c_int), value :: which
    real(c_double), intent(in) :: a(neq, nxmin:nxmax, nymin:nymax, nzmin:nzmax)
    if (which == 0) u = a
    if (which == 1) up = a
    if (which == 2) primit = a
  end subroutine ref_set_array

  ! which: 0 u, 1 up, 2 primit, 3 f, 4 g, 5 h
  subroutine ref_get_array(which, a) bind(C, name='ref_get_array')
    integer(c_int), value :: which
    real(c_double), intent(out) :: a(neq, nxmin:nxmax, nymin:nymax, nzmin:nzmax)
    if (which == 0) a = u
    if (which == 1) a = up
    if (which == 2) a = primit
    if (which == 3) a = f
    if (which == 4) a = g
    if (which == 5) a = h
  end subroutine ref_get_array

  subroutine ref_get_temp(a) bind(C, name='ref_get_temp')
    real(c_double), intent(out) :: a(nxmin:nxmax, nymin:nymax, nzmin:nzmax)
    a = Temp
  end subroutine ref_get_temp

  subroutine ref_boundary(order) bind(C, name='ref_boundary')
    integer(c_int), value :: order
    if (order == 1) call boundaryI()
    if (order == 2) call boundaryII()
  end subroutine ref_boundary

  ! which: 0 calcprim(u, primit), 1 calcprim(up, primit)
  subroutine ref_calcprim(which) bind(C, name='ref_calcprim')
    integer(c_int), value :: which
    if (which == 0) call calcprim(u, primit)
    if (which == 1) call calcprim(up, primit)
  end subroutine ref_calcprim

  subroutine ref_fluxes(choice) bind(C, name='ref_fluxes')
    integer(c_int), value :: choice
    if (riemann_solver == SOLVER_HLLD) call hlldfluxes(choice)
    if (riemann_solver == SOLVER_HLLE) call hllefluxes(choice)
    if (riemann_solver == SOLVER_HLL ) call hllfluxes(choice)
    if (riemann_solver == SOLVER_HLLC) call hllcfluxes(choice)
  end subroutine ref_fluxes

  subroutine ref_get_timestep(iter, n_iter, t, tprint, dt, dump) bind(C, name='ref_get_timestep')
    integer(c_int), value :: iter, n_iter
    real(c_double), value :: t, tprint
    real(c_double), intent(out) :: dt
    integer(c_int), intent(out) :: dump
    logical :: dflag
    real :: tt, tp, dtl
    dflag = .false.
    tt = t
    tp = tprint
    call get_timestep(iter, n_iter, tt, tp, dtl, dflag)
    dt = dtl
    dump = merge(1, 0, dflag)
  end subroutine ref_get_timestep

  ! source(i,j,k,prim,s) (sources.f90:188-220) from the current primit
  subroutine ref_source(i, j, k, prim, sv) bind(C, name='ref_source')
    integer(c_int), value :: i, j, k
    real(c_double), intent(in)  :: prim(neq)
    real(c_double), intent(out) :: sv(neq)
    call source(i, j, k, prim, sv)
  end subroutine ref_source

  ! logical switches of this variant: eight_wave, user_source_terms, bc_user
  subroutine ref_switches(sw) bind(C, name='ref_switches')
    integer(c_int), intent(out) :: sw(3)
    sw = (/ merge(1, 0, eight_wave), merge(1, 0, user_source_terms), merge(1, 0, bc_user) /)
  end subroutine ref_switches

  ! globals::time, which the user hooks read
  subroutine ref_set_time(t) bind(C, name='ref_set_time')
    real(c_double), value :: t
    time = t
  end subroutine ref_set_time

  ! cooling tables as the reference reads them (kind 3: DMC, 4: CHIANTI),
  ! cooltab(2, 41)
  subroutine ref_cooling_table(kind, tab) bind(C, name='ref_cooling_table')
    integer(c_int), value :: kind
    real(c_double), intent(out) :: tab(2, 41)
    if (kind == COOL_DMC) then
      if (.not. allocated(cooltab_dmc)) call init_cooling_dmc()
      tab = cooltab_dmc
    else
      if (.not. allocated(cooltab_chianti)) call init_cooling_chianti()
      tab = cooltab_chianti
    end if
  end subroutine ref_cooling_table

  ! one cooling pass (coolingdmc / coolingchi) on globals u, primit with
  ! dt_CFL = dt
  subroutine ref_cooling(kind, dt) bind(C, name='ref_cooling')
    integer(c_int), value :: kind
    real(c_double), value :: dt
    dt_CFL = dt
    if (kind == COOL_H) then
#ifdef PASSIVES
      call coolingh()
#endif
    else if (kind == COOL_DMC) then
      if (.not. allocated(cooltab_dmc)) call init_cooling_dmc()
      call coolingdmc()
    else
      if (.not. allocated(cooltab_chianti)) call init_cooling_chianti()
      call coolingchi()
    end if
  end subroutine ref_cooling

  ! thermal_conduction (thermal_cond.f90:690-768) with dt_CFL = dt, on the
  ! globals as tstep leaves them (the caller ran boundaryI and calcprim)
  subroutine ref_thermal(dt) bind(C, name='ref_thermal')
    real(c_double), value :: dt
    dt_CFL = dt
    call thermal_conduction()
  end subroutine ref_thermal

  subroutine ref_dt_cond(dt) bind(C, name='ref_dt_cond')
    real(c_double), intent(out) :: dt
    real :: d
    call get_dt_cond(d)
    dt = d
  end subroutine ref_dt_cond

  ! th_cond, tc_saturation, Tempsc of this variant
  subroutine ref_conduction_config(kind, sat, tempsc_out) bind(C, name='ref_conduction_config')
    integer(c_int), intent(out) :: kind, sat
    real(c_double), intent(out) :: tempsc_out
    kind = th_cond
    sat = merge(1, 0, tc_saturation)
    tempsc_out = Tempsc
  end subroutine ref_conduction_config

  ! eq_of_state of this variant
  subroutine ref_eos(eos) bind(C, name='ref_eos')
    integer(c_int), intent(out) :: eos
    eos = eq_of_state
  end subroutine ref_eos

end module ref_kat
