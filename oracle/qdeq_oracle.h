/*
 * qdeq_oracle.h -- CPU FP64 restatement of the reference quasi-dynamic
 * earthquake-cycle solver (XuSun-SUSTech/QDEqCycle_FDM).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * library is compared against; it is never linked into, loaded by or called
 * from the product path.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may use it.
 *
 * Every function cites the reference file:line it restates.  Layouts follow
 * the reference exactly:
 *   - grid arrays are nx*nz row-major, x fastest, one ghost layer
 *     (idx = i + k*nx, struct.h:8, odeSolver.cpp:24-37);
 *   - the multigrid works on a compact ni*nk array (idx = i' + k'*ni,
 *     odeSolver.cpp:669-677) with level dims N=(N-1)/2+1 (odeSolver.cpp:1394);
 *   - mu_equi is the doubled (2nx-1)*(2nz-1) array (odeSolver.cpp:47,90);
 *   - fault arrays have length nz and are valid for k in [1, nk]
 *     (Alloc.cpp:3-40, fault.cpp:94).
 */
#ifndef QDEQ_ORACLE_H
#define QDEQ_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_LEVELS 12

/* Multigrid stop rule.  ORC_TOL_ABS is the reference rule ||r||_2 <= tol
 * (odeSolver.cpp:689-696, tol = 1e-9).  ORC_TOL_REL stops at
 * ||r||_2 <= tol * ||r(u_interior = 0)||_2 (boundary-data residual). */
enum { ORC_TOL_ABS = 0, ORC_TOL_REL = 1 };

/* status codes */
enum {
  ORC_OK = 0,
  ORC_ERR_ARG = 1,
  ORC_ERR_NAN_FRICTION = 2,   /* fault.cpp:211-219 (reference calls exit(0)) */
  ORC_ERR_MG_NOT_CONVERGED = 3
};

typedef struct orc_params {
  int ni, nk;          /* interior nodes in x (fault-normal) and z (depth) */
  double dh;           /* grid spacing, m */
  int n_levels;        /* multigrid levels (odeSolver.cpp:3) */
  int n_cycles;        /* RBGS sweeps per smoothing (odeSolver.cpp:4) */
  double relax;        /* RBGS relaxation (odeSolver.cpp:2) */
  int tol_mode;        /* ORC_TOL_ABS / ORC_TOL_REL */
  double tol;          /* 1e-9 (ABS, odeSolver.cpp:689) or relative tolerance */
  int max_vcycles;     /* V-cycle cap per solve (the reference has none) */
  double ini_dt, mindt, maxdt, totTol, tmax;   /* SetParams.cpp:5-18 */
  double f0, V0, L, Vini;                      /* fault.cpp:85-89 */
  /* cycle shape: 0 = the reference's (n_cycles sweeps everywhere, last fine
   * row uncorrected); 1 = the engine's fast cycle (n_pre / n_post /
   * n_coarse sweeps, every fine row corrected) */
  int cycle, n_pre, n_post, n_coarse;
} orc_params;

/* ---- medium (odeSolver.cpp:337-589) ---- */
void orc_mu_equi(int ni, int nk, const double *mu, double *mu_equi);

/* OpenMP threads of the multigrid loops (n > 0 sets; returns the count in
 * use; 1 in a build without OpenMP).  Results do not depend on it. */
int orc_set_threads(int n);

/* ---- multigrid building blocks (odeSolver.cpp:741-1147) ---- */
void orc_level_dims(int ni, int nk, double dh, int n_levels, int *N, int *K,
                    int *fin_coa, double *DH);
void orc_rbgs(double *u, const double *f, const double *mu_equi, int ni,
              int nk, double dh, int level, int phase, double relax);
void orc_residual(double *r, const double *u, const double *f,
                  const double *mu_equi, int ni, int nk, double dh, int level);
void orc_restrict(double *r2, const double *r, double *e2, int ni, int nk,
                  int level);
void orc_prolong(double *u, const double *e2, int ni, int nk, int level);
void orc_prolong_full(double *u, const double *e2, int ni, int nk, int level,
                      int full_top);
double orc_nrm2(const double *x, long n);

/* opaque multigrid solver: owns the level buffers (odeSolver.cpp:6-8,1391-1405) */
typedef struct orc_mg orc_mg;
orc_mg *orc_mg_create(int ni, int nk, double dh, int n_levels, int n_cycles,
                      double relax, const double *mu_equi);
void orc_mg_destroy(orc_mg *mg);
/* sweep counts and top-row correction of the cycle (default: the reference's) */
void orc_mg_set_cycle(orc_mg *mg, int n_pre, int n_post, int n_coarse,
                      int full_top);
/* one V-cycle on the compact level-0 arrays (odeSolver.cpp:1101-1147) */
void orc_mg_vcycle(orc_mg *mg, double *u, const double *f);
/* ode_solve = multigrid + extend_crew on the nx*nz array U
 * (odeSolver.cpp:605-609, 652-720, 612-650) */
int orc_ode_solve(orc_mg *mg, double *U, int tol_mode, double tol,
                  int max_vcycles, int *vcycles, double *res0, double *res);
void orc_extend_crew(double *U, int ni, int nk);

/* ---- fault physics (fault.cpp:139-232, odeSolver.cpp:1150-1206) ---- */
void orc_set_boundary(double *U, int ni, int nk, const double *slip,
                      const double *slip_shift, double Vini, double t,
                      double ratio, double dt);
void orc_eval_stress(const double *U, const double *mu, const double *tau_n0,
                     int ni, int nk, double dh, double *Ts, double *Tn);
int orc_nr_search(int nk, const double *Ts, const double *Tn,
                  const double *psi, const double *a, const double *eta,
                  double V0, double *V, int *iters);
void orc_cal_dpsi(int nk, const double *psi, const double *V, const double *b,
                  double f0, double V0, double L, double *dpsi);
double orc_compute_error(int nk, const double *slip3, const double *psi3,
                         const double *slip4, const double *psi4);
double orc_compute_step_size(int it, double dt, double mindt, double maxdt,
                             double totErr, double totTol, const double *errA);

/* Elementary functions of NRsearch / cal_dpsi: 0 = the engine's
 * host/device-identical algorithms (default), 1 = glibc. */
void orc_set_libm(int use_glibc);
/* fn: 0 exp, 1 sinh, 2 cosh, 3 log, 4 log1p, 5 asinh (engine algorithms) */
double orc_elem(int fn, double x);

/* ---- RK43 driver (odeSolver.cpp:1277-1674) ---- */
/* Butcher tableau of odeSolver.cpp:1297-1334: c[6], b[6], hb[6], a[6][6]. */
void orc_tableau(double *out54);
typedef struct orc_sim orc_sim;
/* U0: nx*nz initial displacement with ghosts filled by extend_crew
 * (eq_cycle.cpp:25); a,b,eta,tau_n0: nz; mu: nx*nz nodal; mu_equi doubled. */
orc_sim *orc_sim_create(const orc_params *p, const double *mu,
                        const double *mu_equi, const double *a,
                        const double *b, const double *eta,
                        const double *tau_n0, const double *U0);
void orc_sim_destroy(orc_sim *s);
int orc_sim_init(orc_sim *s);          /* odeSolver.cpp:1428-1437 */
int orc_sim_set_state(orc_sim *s, double t, double dt, int it, double errA0,
                      double errA1, const double *U, const double *slip,
                      const double *psi);
int orc_sim_step(orc_sim *s);          /* one accepted step, 1462-1633 */
/* Test hooks of the RK43 checks: the step size of the next attempt, the
 * last attempt's stage rates / states (stages 2..6 at offset (i-1)*nz) and
 * y3 / y4, and the log of the last stage calls (t, dt, ratio, U at row 1 of
 * the remote and of the fault column after setBoundary; 5 doubles each,
 * oldest first, at most 8; returns the count). */
void orc_sim_set_dt(orc_sim *s, double dt);
void orc_sim_last_attempt(const orc_sim *s, double *fV, double *fd, double *ks,
                          double *kp, double *y3s, double *y3p, double *y4s,
                          double *y4p);
int orc_sim_stage_log(const orc_sim *s, double *out);
/* scalars: t, dt, it, errA0, errA1, totErr, attempts, vcycles(total),
 * solves(total), V_max, Ts_max */
void orc_sim_scalars(const orc_sim *s, double *out11);
void orc_sim_arrays(const orc_sim *s, double *U, double *slip, double *psi,
                    double *V, double *dpsi, double *Ts);

#ifdef __cplusplus
}
#endif
#endif
