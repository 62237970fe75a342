/*
 * qdeq.h -- C ABI of the MI355X-native quasi-dynamic earthquake-cycle engine
 * (libqdeq.so, built from qdeqcycle_fdm_amd/csrc for gfx950).
 *
 * The entry points replace the reference's hot-path call surface
 * (XuSun-SUSTech/QDEqCycle_FDM functions.h:11-45): adaptive RK43 stepping that
 * calls the antiplane elliptic multigrid solve and then the fault traction /
 * rate-and-state update.  Each replaced function is cited next to its entry
 * point.
 *
 * Conventions (the reference's, so callers can hand over their arrays):
 *   - FP64 everywhere at the boundary ("Real = double", struct.h:4);
 *   - grid arrays are nx*nz, x fastest, one ghost layer: idx = i + k*nx with
 *     nx = ni+2, nz = nk+2 (odeSolver.cpp:24-37); column 1 is the fault,
 *     column ni the remote edge; row 1 the bottom, row nk the free surface;
 *   - mu_equi is the doubled (2nx-1)*(2nz-1) equivalent-medium array
 *     (odeSolver.cpp:47,90);
 *   - fault arrays have length nz, valid on rows 1..nk (Alloc.cpp:3-40);
 *   - units: m, GPa, MPa, MPa*s/m, m/s, s (fault.cpp:153, coord.cpp:18).
 * Differences from the reference:
 *   - every entry point returns a qdeq_status instead of calling exit()
 *     (helper_cuda.h:540-555, fault.cpp:218);
 *   - array arguments carry a pointer-space flag (QDEQ_HOST / QDEQ_DEVICE);
 *     every array of one call lives in that space; device arrays must be on
 *     the context's device, and work the caller queued on the null stream
 *     (stream 0) is complete before the call reads them;
 *   - all solver state lives in a context (the reference keeps it in
 *     file-scope globals, odeSolver.cpp:6-12), so several contexts may coexist;
 *     one context must not be used from two host threads at once.
 */
#ifndef QDEQ_H
#define QDEQ_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum qdeq_status {
  QDEQ_OK = 0,
  QDEQ_ERR_ARG = 1,               /* invalid argument or size */
  QDEQ_ERR_HIP = 2,               /* HIP runtime error */
  QDEQ_ERR_NAN_FRICTION = 3,      /* NR produced NaN twice (fault.cpp:211-219) */
  QDEQ_ERR_MG_NOT_CONVERGED = 4,  /* V-cycle cap reached before tolerance */
  QDEQ_ERR_STATE = 5,             /* call out of order (e.g. step before setup) */
  QDEQ_ERR_NOMEM = 6
} qdeq_status;

enum { QDEQ_HOST = 0, QDEQ_DEVICE = 1 };
/* elliptic stop rule: reference ||r||_2 <= tol (odeSolver.cpp:689-696), or
 * ||r||_2 <= tol * ||r(u_interior = 0)||_2 (boundary-data residual) */
enum { QDEQ_TOL_ABS = 0, QDEQ_TOL_REL = 1 };
/* V-cycle arithmetic: plain FP64, or FP32 V-cycles inside FP64 iterative
 * refinement (FP64 residual and update) */
enum { QDEQ_PREC_FP64 = 0, QDEQ_PREC_FP32_IR = 1 };
/* V-cycle shape: the reference's exactly (odeSolver.cpp:1101-1147, incl. the
 * uncorrected top row), or the same cycle with the sweep counts of the
 * config and a bilinear correction of the top row as well */
enum { QDEQ_CYCLE_REFERENCE = 0, QDEQ_CYCLE_FAST = 1 };

typedef struct qdeq_config {
  int ni, nk;        /* interior nodes in x and z (SetParams.cpp:21-22) */
  double dh;         /* grid spacing in m (SetParams.cpp:23) */
  int n_levels;      /* multigrid levels (odeSolver.cpp:3) */
  int n_pre, n_post, n_coarse; /* sweeps (odeSolver.cpp:4: 32 each); with the
                                  fast cycle n_coarse = 0 solves the coarsest
                                  level exactly (at most 1024 unknowns) */
  double relax;      /* RBGS relaxation (odeSolver.cpp:2) */
  int cycle;         /* QDEQ_CYCLE_* */
  int precision;     /* QDEQ_PREC_* */
  int tol_mode;      /* QDEQ_TOL_* */
  double tol;        /* 1e-9 absolute in the reference (odeSolver.cpp:689) */
  int max_vcycles;   /* cap per solve (> 0); the reference has none */
  double ini_dt, mindt, maxdt, totTol, tmax;   /* SetParams.cpp:5-18 */
  double f0, V0, L, Vini;                      /* fault.cpp:85-89 */
  int device;        /* HIP device ordinal */
} qdeq_config;

typedef struct qdeq_ctx qdeq_ctx;

/* Per accepted step record (the reference prints it, odeSolver.cpp:1636). */
typedef struct qdeq_step_info {
  double t, dt;              /* time after the step, next step size */
  double totErr;             /* embedded error of the accepted step */
  double errA0, errA1;       /* controller history (odeSolver.cpp:1352) */
  double V_max, Ts_max;      /* maxval (fault.cpp:22-52) */
  int it;                    /* accepted steps so far */
  int attempts;              /* attempts of the last step */
  long long vcycles;         /* V-cycles (FP64 ones, or FP32 ones in IR) */
  long long solves;          /* elliptic solves */
  long long ir_sweeps;       /* refinement iterations (FP32_IR only) */
} qdeq_step_info;

/* Fills the reference's settings (SetParams.cpp, odeSolver.cpp:2-4, 689). */
void qdeq_default_config(qdeq_config *cfg);

/* Creates a context on cfg->device: level hierarchy (odeSolver.cpp:1384-1420),
 * operator coefficients and device buffers.  mu: nodal shear modulus, nx*nz
 * host array (MEDIUM.mu).  mu_equi: doubled equivalent-medium host array, or
 * NULL to build it on the device from mu (SetMedium, odeSolver.cpp:337-589). */
int qdeq_create(const qdeq_config *cfg, const double *mu,
                const double *mu_equi, qdeq_ctx **out);
int qdeq_destroy(qdeq_ctx *ctx);
/* Execution shape of the fast cycle (no reference counterpart): tiled != 0
 * runs it on the temporally blocked LDS kernels with interior tiles of
 * tile_x by tile_y points (even sizes, limited by the 160 KiB LDS at the
 * configured sweep counts); tiled == 0 uses one launch per half-sweep.
 * Results are identical either way. */
int qdeq_set_tiles(qdeq_ctx *ctx, int tiled, int tile_x, int tile_y);
/* 1 if solves of ctx currently run the fast cycle on the tiled kernels. */
int qdeq_uses_tiled_cycle(const qdeq_ctx *ctx);
/* Tile-shape autotuner (no reference counterpart): for every level run by
 * the tile kernels, times `reps` launches of the down (smoothing +
 * restriction) and the up (prolongation + smoothing) kernel for each
 * candidate shape (interior tile_x x tile_y, thread block 64 x rows) that
 * fits the LDS, and keeps per level and kernel the first candidate (in a
 * fixed order) within 3% of the fastest.  Solutions are
 * identical for every shape.  The level-0 displacement is preserved, so it
 * may be called at any time; qdeq_set_tiles discards the tuned shapes. */
int qdeq_autotune_tiles(qdeq_ctx *ctx, int reps);
/* Shapes in use: *nlevels tile levels, shapes[6*l .. 6*l+5] = down tile_x,
 * tile_y, thread rows, up tile_x, tile_y, thread rows (6*12 ints). */
int qdeq_get_tile_shapes(const qdeq_ctx *ctx, int *nlevels, int *shapes);
/* Pins the shapes (same layout; nlevels must equal the tile level count),
 * e.g. to reuse a tuning; rejected if any does not fit the kernels. */
int qdeq_set_tile_shapes(qdeq_ctx *ctx, int nlevels, const int *shapes);
/* Average duration (microseconds) of each kernel of the tiled cycle, each
 * launched `reps` times back to back on the context's buffers (the level-0
 * displacement is preserved): *n = 2*levels+1 entries, the down kernels of
 * the tile levels from the finest, the tail, the up kernels back to level 0
 * (us: up to 2*12+1 doubles). */
int qdeq_kernel_breakdown(qdeq_ctx *ctx, int reps, int *n, double *us);

/* Measurement hooks (no reference counterpart).  qdeq_set_kernel_timing
 * brackets one kernel of the tiled fast cycle with HIP events on the
 * context's stream: every 8th V-cycle is launched kernel by kernel (also
 * where V-cycles are otherwise graph replays; same kernels, same arguments)
 * with the events recorded right before and after that kernel;
 * qdeq_get_kernel_timing reads the accumulated duration and the count of
 * timed launches (a sampled V-cycle that the stop test skipped is not
 * counted).  qdeq_copy_probe times `reps` launches of a plain device copy
 * of n_doubles doubles (bandwidth and counter calibration). */
enum { QDEQ_TIME_OFF = 0, QDEQ_TIME_UP0 = 1, QDEQ_TIME_TAIL = 2,
       QDEQ_TIME_DOWN1 = 3, QDEQ_TIME_FAULT = 4 };
/* QDEQ_TIME_FAULT times the fault update after each stage solve (traction,
 * Newton solve for V, state rate), every 8th one, with events around it. */
int qdeq_set_kernel_timing(qdeq_ctx *ctx, int which);
/* hipDeviceSynchronize on `device` (a barrier for host-side timing). */
int qdeq_device_synchronize(int device);
int qdeq_get_kernel_timing(qdeq_ctx *ctx, double *total_ms, long long *calls);
int qdeq_copy_probe(qdeq_ctx *ctx, long long n_doubles, int reps, double *avg_ms);

/* Message of the last failing call on ctx ("" if none). */
const char *qdeq_last_error(const qdeq_ctx *ctx);
/* Message of the last failing qdeq_create / context-free call. */
const char *qdeq_global_error(void);
/* HIP runtime version (HIP_VERSION encoding: major*1e7 + minor*1e5 + patch)
 * the library was built against, and the one the process runs (no
 * reference counterpart).  They differ when another library loaded first
 * brought its own HIP runtime under the same soname; callers should refuse
 * to run then.  running may be NULL. */
int qdeq_hip_runtime_version(int *built, int *running);

/* Writes the doubled equivalent-medium array the context uses (built on the
 * device when qdeq_create got NULL).  where: QDEQ_HOST / QDEQ_DEVICE. */
int qdeq_get_mu_equi(qdeq_ctx *ctx, double *mu_equi, int where);
/* Context-free device build of SetMedium's mu_equi (odeSolver.cpp:337-589)
 * from the nodal nx*nz host array mu, on HIP device `device`. */
int qdeq_build_mu_equi(int ni, int nk, const double *mu, double *mu_equi,
                       int device);

/* The same build that allocates nothing (no reference counterpart; for
 * callers that cannot allocate device memory, such as loops that rebuild
 * the medium of many ensemble members without touching the allocator): mu
 * (nodal nx*nz, ghost entries ignored) and mu_equi ((2nx-1)*(2nz-1)) are
 * device arrays on `device`; the work runs on the null stream and the call
 * returns when it is done.  The interpolation weight tables live in a
 * fixed 76,800-byte device array that each device keeps until the process
 * ends; every call recomputes them, and calls that overlap on one device
 * race on it.  Grids of 3..65535 nodes per side; the result equals
 * qdeq_build_mu_equi's bit for bit. */
int qdeq_build_mu_equi_noalloc(int ni, int nk, const double *mu,
                               double *mu_equi, int device);

/* ---- drop-in replacements of the reference's hot-path functions ---- */

/* ode_solve (odeSolver.cpp:605-609) = multigrid (652-720) + extend_crew
 * (612-650): solves div(mu grad u) = 0 for the interior of U with U's fault
 * and remote columns as Dirichlet data, starting from U (warm start), then
 * extrapolates the ghost layer.  U: nx*nz.  Outputs (optional, may be NULL):
 * V-cycles done, boundary-data (REL) or initial (ABS) residual norm, final
 * residual norm. */
int qdeq_ode_solve(qdeq_ctx *ctx, double *U, int where, int *vcycles,
                   double *res0, double *res);

/* setBoundary (odeSolver.cpp:1150-1188): fault column = slip/2, remote column
 * = Vini*(t + ratio*dt)/2 + slip_shift, mirrored ghost rows.  U: nx*nz;
 * slip, slip_shift: nz. */
int qdeq_set_boundary(qdeq_ctx *ctx, double *U, const double *slip,
                      const double *slip_shift, double t, double ratio,
                      double dt, int where);

/* eval_stress_on_fault + partial1st (fault.cpp:139-164): Ts = mu*du/dx*1e3
 * at the fault column (one-sided 4-point stencil), Tn = tau_n0. */
int qdeq_eval_stress(qdeq_ctx *ctx, const double *U, const double *mu,
                     const double *tau_n0, double *Ts, double *Tn, int where);

/* NRsearch (fault.cpp:169-232): per node, Newton on the asinh variable for
 * the slip rate with radiation damping.  V holds the previous slip rate on
 * entry (the Newton start) and the new one on exit.  Returns
 * QDEQ_ERR_NAN_FRICTION where the reference would exit(0). */
int qdeq_nr_search(qdeq_ctx *ctx, const double *Ts, const double *Tn,
                   const double *psi, const double *a, const double *eta,
                   double *V, int where);

/* cal_dpsi (odeSolver.cpp:1190-1206): aging-law state rate. */
int qdeq_cal_dpsi(qdeq_ctx *ctx, const double *psi, const double *V,
                  const double *b, double *dpsi, int where);

/* computeError (odeSolver.cpp:1211-1238) on host arrays of length nk+2. */
double qdeq_compute_error(int nk, const double *slip3, const double *psi3,
                          const double *slip4, const double *psi4);
/* computeStepSize (odeSolver.cpp:1241-1271). errA: 2 doubles. */
double qdeq_compute_step_size(int it, double dt, double mindt, double maxdt,
                              double totErr, double totTol,
                              const double *errA);

/* ---- RK43 driver (odeSolver.cpp:1277-1674), device resident ---- */

/* fault_init (fault.cpp:116-133) + the t=0 stage (odeSolver.cpp:1428-1437).
 * a, b, eta, tau_n0: nz host arrays; U0: nx*nz host initial displacement
 * (ghosts as extend_crew leaves them).  Slip = 2 U0(fault), SlipShift =
 * U0(remote), psi = f0 + b ln(V0/Vini). */
int qdeq_rk43_setup(qdeq_ctx *ctx, const double *a, const double *b,
                    const double *eta, const double *tau_n0, const double *U0);
/* Advances nsteps accepted steps (odeSolver.cpp:1462-1633), fewer if t
 * reaches tmax (the last step is shortened to end there).  If info is not
 * NULL it receives the record of the last step. */
int qdeq_rk43_step(qdeq_ctx *ctx, int nsteps, qdeq_step_info *info);
/* Ensemble stepping (no reference counterpart): advances nsteps accepted
 * steps of each of n independent contexts (any grids and settings, each
 * after qdeq_rk43_setup, each on its own stream) from one host thread.  The
 * driver is a resumable state machine: it never waits on one member while
 * another can launch work, so small grids' latency-bound kernels overlap.
 * Each member runs the same launches in the same order as qdeq_rk43_step,
 * so its results are bit-identical to running it alone.  status[i] (may be
 * NULL) receives each member's status; the return value is QDEQ_OK or the
 * last failure. */
int qdeq_ensemble_step(qdeq_ctx **ctxs, int n, int nsteps, int *status);
/* ---- output records and event catalogue ---- */

/* The reference's fault output (odeSolver.cpp:1442-1457, 1638-1652): writes
 * dir/output_fault_0.bin now (t, V, Slip, Ts, psi, U, Tn, a, b, dt after the
 * t = 0 stage, which runs first if needed) and, after every `every`-th
 * accepted step (it_save_jump, SetParams.cpp:27), dir/output_fault_<k>.bin
 * with t, V, Slip, Ts, psi, dt (FP64, nz per array, native byte order).
 * Records are copied into pinned host slots on the context's stream and
 * written once complete; qdeq_output_close (and qdeq_destroy) flush them. */
int qdeq_output_open(qdeq_ctx *ctx, const char *dir, int every);
int qdeq_output_close(qdeq_ctx *ctx);

/* Event catalogue (no reference counterpart): after every accepted step a
 * kernel tracks max V; an event opens at the first step with max V above
 * v_threshold and closes at the first step back at or below it.  Times are
 * those after the step; nodes are fault rows 1..nk; slip_max is the largest
 * slip accumulated on any node since the onset; an open event has
 * t_end = -1, it_end = -1. */
typedef struct qdeq_event {
  double t_onset, t_end, t_peak, v_peak, slip_max;
  int it_onset, it_end, node_onset, node_peak;
} qdeq_event;
int qdeq_events_enable(qdeq_ctx *ctx, double v_threshold, int capacity);
/* Copies up to max entries; *n = events recorded (may exceed capacity). */
int qdeq_events_get(qdeq_ctx *ctx, qdeq_event *out, int max, int *n);

/* Projection start of the RK43 stage solves (no reference counterpart: the
 * reference starts every solve from U as the previous solve left it,
 * odeSolver.cpp:652-720, which stays the fallback).  The last `slots`
 * solutions (1..8; 0 disables) are kept with their fault columns; each
 * solve is offered the combination of them whose fault data best fit the
 * new fault data in the least-squares sense, plus the exact response to the
 * remote data, and uses it when its residual norm is finite and below the
 * fallback's.  Enabling runs one solve (the unit remote response) and
 * allocates slots + 1 level-0 fields; exclusive with qdeq_attach_guess. */
int qdeq_projection_enable(qdeq_ctx *ctx, int slots);
/* The least-squares fit of the projection start on host arrays, by the
 * kernel the RK43 driver uses (for checking it): g_new (N = fault nodes of
 * the context), g_old (cnt x N, newest first, 2 <= cnt <= 8).  coef (cnt + 2
 * doubles): coef[0] = 0, coef[i] = c_i for i = 1..cnt-1 (0 for a column
 * left out), coef[cnt] = ||misfit|| / ||g_new - g_old[0]||, coef[cnt + 1] =
 * columns kept. */
int qdeq_projection_fit(qdeq_ctx *ctx, const double *g_new,
                        const double *g_old, int cnt, double *coef);
/* Fault injection for tests (no reference counterpart): every later
 * projection start is U_0 + factor (start - U_0), U_0 the newest stored
 * solution (1 = off).  A factor such as 1e6 or NaN makes the start worse
 * than the previous solution, so the solves fall back to it. */
int qdeq_projection_inject(qdeq_ctx *ctx, double factor);
/* Starts offered and used (shared with the learned-guess counters), and the
 * last fit's ||residual|| / ||fault data change|| and column count. */
int qdeq_projection_stats(qdeq_ctx *ctx, long long *offered,
                          long long *used, double *fit_ratio, int *columns);

/* Solves launched ahead by qdeq_rk43_step (no reference counterpart): with
 * one simulation launching kernel by kernel, each stage's fault update,
 * start and first V-cycles are queued behind the previous stage's solve and
 * run only if that solve converged with the V-cycles already queued; else
 * they abandon themselves and the solve goes on.  The same holds across
 * steps (the next step's stage 2 behind an accepted step's final stage 1)
 * when the next step runs within the same call and neither per-step outputs
 * nor events are on.  Results are the same either way (QDEQ_SPECULATE=0 in
 * the environment at qdeq_create turns it off).  Counts of such solves taken
 * over and abandoned, and of those taken, the ones across a step (any
 * pointer may be NULL). */
int qdeq_launch_ahead_stats(qdeq_ctx *ctx, long long *taken, long long *abandoned,
                            long long *taken_next_step);

/* Stop tests folded into the next V-cycle so far (no reference
 * counterpart): within a batch of V-cycles launched kernel by kernel (one
 * simulation, FP64 tiled cycle), the residual-norm test of every V-cycle
 * but the batch's last runs in the next V-cycle's level-0 down kernel
 * instead of a kernel of its own; every workgroup sums the norm partials in
 * the stop-test kernel's order, so the decisions and results are the same
 * (QDEQ_FOLD_STOP=0 in the environment at qdeq_create turns it off). */
int qdeq_folded_stop_tests(qdeq_ctx *ctx, long long *folded);

/* Copies the current state to host arrays (any may be NULL): U (nx*nz, with
 * ghosts as ode_solve leaves them), slip, psi, V, dpsi, Ts (nz). */
int qdeq_rk43_get_state(qdeq_ctx *ctx, qdeq_step_info *info, double *U,
                        double *slip, double *psi, double *V, double *dpsi,
                        double *Ts);

#ifdef __cplusplus
}
#endif
#endif /* QDEQ_H */
