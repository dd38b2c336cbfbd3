/*
 * guacho_oracle.h -- CPU restatement of Guacho-3D's Godunov MHD step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle for the HIP
 * product path: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.  It is never linked into, or called from,
 * the product library (guacho_amd/libguacho_amd.so).
 *
 * Layout follows the reference exactly: arrays are Fortran column-major
 * u(neq, nxmin:nxmax, nymin:nymax, nzmin:nzmax) with neq = 8 fastest and
 * nghost = 2 (src/globals.f90:33-38, src/init.f90:145-158).  Indices in
 * the API below are the reference's 1-based cell indices; interior cells
 * are 1..nx, ghosts are -1,0 and nx+1,nx+2.
 */
#ifndef GUACHO_ORACLE_H
#define GUACHO_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Named constants, values as src/constants.f90:56-93. */
enum { ORC_SOLVER_HLL = 1, ORC_SOLVER_HLLC = 2, ORC_SOLVER_HLLE = 3, ORC_SOLVER_HLLD = 4 };
enum { ORC_BC_OUTFLOW = 1, ORC_BC_CLOSED = 2, ORC_BC_PERIODIC = 3, ORC_BC_OTHER = 4 };
enum {
  ORC_LIMITER_NO_AVERAGE = -1, ORC_LIMITER_NO_LIMIT = 0, ORC_LIMITER_MINMOD = 1,
  ORC_LIMITER_VAN_LEER = 2, ORC_LIMITER_VAN_ALBADA = 3, ORC_LIMITER_UMIST = 4,
  ORC_LIMITER_WOODWARD = 5, ORC_LIMITER_SUPERBEE = 6
};

typedef struct orc_params {
  int nx, ny, nz;         /* interior cells */
  double dx, dy, dz;
  double cv, gamma;       /* gamma = (cv+1)/cv as OT/parameters.f90:158 */
  double eta;             /* artificial viscosity (hydro_solver.f90:57) */
  int solver;             /* ORC_SOLVER_* */
  int limiter;            /* ORC_LIMITER_* */
  int bc[6];              /* left, right, bottom, top, out, in */
  int mhd, pmhd;          /* physics flags (parameters.f90:48-49) */
  int flux_cd;            /* enable_flux_cd (parameters.f90:63) */
  int eight_wave;         /* Powell 8-wave source terms (parameters.f90:61) */
  int user_src;           /* the oracle's user source (oracle/ref/user_mod.f90) */
  int user_bc;            /* bc_user: the oracle's impose_user_bc after boundaryI/II */
  int nthreads;           /* OpenMP threads for the sweeps (<=0: runtime default) */
  int eos;                /* eq_of_state: 2 EOS_SINGLE_SPECIE, 3 EOS_H_RATE, else adiabatic */
  double tempsc;          /* Tempsc of the EOS_SINGLE_SPECIE / EOS_H_RATE temperature */
  int npas;               /* passive scalars after the 8 MHD variables (0..2) */
} orc_params;

typedef struct orc_state orc_state;

/* ---- whole-domain API (Fortran layout arrays incl. ghosts) ---------- */
int  orc_create(const orc_params *p, orc_state **out);
void orc_destroy(orc_state *s);
long orc_cells(const orc_state *s);           /* (nx+4)(ny+4)(nz+4) */
void orc_set_u(orc_state *s, const double *u);
void orc_get_u(const orc_state *s, double *u);
void orc_set_up(orc_state *s, const double *up);
void orc_get_up(const orc_state *s, double *up);
void orc_set_primit(orc_state *s, const double *primit);
void orc_get_primit(const orc_state *s, double *primit);
void orc_get_temp(const orc_state *s, double *temp);
void orc_get_flux(const orc_state *s, int axis, double *fgh);   /* axis 0:f 1:g 2:h */
void orc_get_e(const orc_state *s, double *e);                  /* e(3, ...) */

void orc_set_time(orc_state *s, double time);   /* globals::time, seen by the user hook */
/* Table cooling after each step (kind 3: coolingdmc, cooling_dmc.f90:115-156;
 * kind 4: coolingchi, cooling_chi.f90), cooltab(2, 41) as the reference reads
 * it; kind 0 turns it off.  tempsc, psc, tsc: Tempsc, Psc, tsc of the
 * parameters module. */
void orc_set_cooling(orc_state *s, int kind, const double *cooltab, double tempsc, double psc, double tsc);
/* one cooling pass on u and primit (interior) with dt_CFL = dt */
void orc_cooling(orc_state *s, double dt);
/* Thermal conduction after each step (thermal_cond.f90:690-768): kind 1
 * isotropic, 2 anisotropic (0 off); sat: tc_saturation; the parameters
 * module's Tempsc, rsc, rhosc, mu, Psc, tsc, vsc2, bsc. */
void orc_set_conduction(orc_state *s, int kind, int sat, double tempsc, double rsc, double rhosc, double mu,
                        double psc, double tsc, double vsc2, double bsc);
/* thermal_conduction with dt_CFL = dt on the current u, primit (primit and
 * its temperature as tstep leaves them); returns the sub-step count */
int orc_thermal(orc_state *s, double dt);
double orc_dt_cond(orc_state *s);   /* get_dt_cond on the current primit */
void orc_boundaryI(orc_state *s);
void orc_boundaryII(orc_state *s);
void orc_calcprim_u(orc_state *s);
void orc_calcprim_up(orc_state *s);
int  orc_fluxes(orc_state *s, int choice);    /* 0 ok, 1 solver fall-through */
void orc_get_efield(orc_state *s);
void orc_step(orc_state *s, double dt);
void orc_viscous_copy(orc_state *s);
int  orc_tstep(orc_state *s, double dt);      /* one full tstep() */
double orc_cfl_dtp(const orc_state *s);       /* min loop of get_timestep */
/* source(i,j,k,prim,s) of sources.f90:188-220: s = 0, the oracle's user
 * source when user_src, then divbcorr_8w_source from the state's primit
 * when eight_wave */
void orc_source(const orc_state *s, int i, int j, int k, const double *prim, double *sv);
void orc_divb(const orc_state *s, double *divb);   /* divB(nx,ny,nz) of u */

/* get_timestep host logic, hydro_core.f90:677-694 */
void orc_timestep_finish(int current_iter, int n_iter, double cfl, double dtp,
                         double current_time, double tprint,
                         double *dt, int *dump_flag);

/* ---- per-cell / per-interface functions (8-vectors) ----------------- */
void orc_u2prim(const orc_params *p, const double *uu, double *prim, double *T);
void orc_prim2u(const orc_params *p, const double *prim, double *uu);
void orc_prim2f(const orc_params *p, const double *prim, double *ff);
void orc_swapy(double *var);
void orc_swapz(double *var);
double orc_csound(const orc_params *p, double pr, double d);
void orc_cfast(const orc_params *p, double pr, double d, double bx, double by, double bz,
               double *cfx, double *cfy, double *cfz);
double orc_cfastX(const orc_params *p, const double *prim);
void orc_limiter(const orc_params *p, const double *pll, double *pl, double *pr, const double *prr);
int  orc_riemann(const orc_params *p, const double *priml, const double *primr, double *ff);
int  orc_hlld(const orc_params *p, const double *priml, const double *primr, double *ff);
int  orc_hlle(const orc_params *p, const double *priml, const double *primr, double *ff);
int  orc_hll(const orc_params *p, const double *priml, const double *primr, double *ff);
int  orc_hllc(const orc_params *p, const double *priml, const double *primr, double *ff);

#ifdef __cplusplus
}
#endif
#endif
