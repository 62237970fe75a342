// qdeq_internal.h -- shared declarations of the HIP engine (not part of the
// public C ABI, which is include/qdeq.h).
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <map>
#include <string>
#include <vector>

#include "qdeq.h"
#include "qdeq_guess.h"

#define QDEQ_MAX_LEVELS 12

namespace qdeq {

// One multigrid level, compact row-major storage with a padded row pitch P
// (elements).  Row j is fault-parallel depth row j (0 = bottom), column k is
// the distance from the fault (0 = fault, K-1 = remote edge).  The level
// hierarchy is N <- (N-1)/2+1, DH <- 2 DH (odeSolver.cpp:1391-1405).
struct LevelDims {
  int N, K, P;   // rows, columns, row pitch in elements
  int fc;        // 2^level: fine nodes per coarse spacing
  double DH;     // grid spacing of the level
};

// Execution shape of one tile kernel launch: interior tile tx x ty points,
// thread block 64 x tyt.
struct TileShape {
  int tx, ty, tyt;
};

// Device buffers of one level.  u: level 0 = displacement (FP64 path) or the
// IR correction (FP32 path); levels >= 1 = coarse correction.  f: right-hand
// side (level 0 FP64 = zero, never read).  r: residual scratch.  cx: west-face
// coefficient of point (j,k) (east face = cx[k+1]); cz: south-face coefficient
// of row j, N+1 rows (north face = row j+1).  Boundary columns of u are
// Dirichlet data (level 0) or zero (corrections).
template <typename T>
struct LevelBufs {
  T *u = nullptr, *f = nullptr, *r = nullptr, *cx = nullptr, *cz = nullptr;
};

// RK43 / fault state on the device, all length nz (valid rows 1..nk).
struct FaultBufs {
  double *slip, *psi, *slip_shift, *a, *b, *eta, *tau_n0, *mu_f, *Ts, *Tn;
  double *Ys[7], *Yp[7];   // stage states k1..k6 (slip, psi)
  double *FV[7], *Fd[7];   // stage rates f1..f6 (V, dpsi)
  double *y3s, *y3p, *y4s, *y4p;
  double *base;            // single allocation backing all of the above
};

// Device scalars (one allocation).
enum {
  SC_NORM = 0,    // residual norm of the last check
  SC_NORM0 = 1,   // boundary-data residual norm
  SC_TOTERR = 2,
  SC_VMAX = 3,
  SC_TSMAX = 4,
  SC_NORM_PRED = 5,  // residual norm of the candidate start
  SC_T = 6,          // RK43 time t_n of the current attempt (the stage kernels
  SC_DT = 7,         //   read t and dt here, so captured stages replay)
  MB_SEQ = 8,        // mailbox only: sequence number of the last V-cycle seen
  MB_COUNT = 9,      //   V-cycles done in the current solve
  MB_CODE = 10,      //   outcome of the last V-cycle (k_norm_gate)
  MB_USED = 11,      //   1 if the solve started from the candidate
  MB_START = 12,     //   residual norm of the start used
  SC_COUNT = 16
};

struct OutputRing;  // output.hip

// State of a solve in progress (mg_begin / mg_poll).
struct MgSolve {
  int phase = 0;        // 0 idle, 1 initial norms pending, 2 V-cycles pending
  bool ir = false, cand = false, sample = false;
  // start decided on the device (k_start_gate); its norms not yet read back
  bool dev_start = false, start_read = false;
  // V-cycles launched in this solve, and the ordinal of the timed one (a
  // timed cycle the gate skipped is not counted)
  int launched = 0, sample_ordinal = 0;
  int it = 0;
  double delta = 0, delta0 = 0, r0 = 0, thr = 0;
  long long seq_target = 0;  // sequence number ending the launched batch
  int page = 0;              // mailbox page its kernels write (see mb_dev)
};

// State of the resumable RK43 driver (api.hip).
struct RkMachine {
  int phase = 0;
  int s = 0;               // stage being solved
  bool staged = false;     // stage s's state already launched (with the update)
  int attempts = 0;        // attempts of the current step
  long long steps_left = 0;
  bool final_stage = false;  // stage 1 after an accepted step
  // the next stage's fault update, start and first batch are queued behind
  // stage s's solve (qdeq_ctx::mg_next), to be taken over if it converges
  bool spec = false;
  // for the final stage 1: the next attempt's dt the queued stage 2 used
  double spec_dt = 0.0;
  hipEvent_t ev = nullptr;   // completion of the last host readback
};

}  // namespace qdeq

struct qdeq_ctx {
  qdeq_config cfg;
  int nx, nz, L;
  qdeq::LevelDims lv[QDEQ_MAX_LEVELS];
  qdeq::LevelBufs<double> d[QDEQ_MAX_LEVELS];
  qdeq::LevelBufs<float> s[QDEQ_MAX_LEVELS];
  double *mu_equi = nullptr;   // doubled equivalent-medium array (device)
  double *partial = nullptr;   // reduction partials (2 x NORM_BLOCKS)
  double *scal = nullptr;      // qdeq::SC_* device scalars
  double *scal_host = nullptr; // pinned mirror of scal
  double *mbox_host = nullptr; // host-mapped norm mailbox (SC_* slots)
  double *mbox_dev = nullptr;  // its device address
  int *flags = nullptr;        // [0] NR NaN flag
  int *flags_host = nullptr;
  double *stage_U = nullptr;   // nx*nz staging for drop-in entry points
  double *stage_v[8] = {};     // nz staging vectors
  qdeq::FaultBufs fb{};
  hipStream_t stream = nullptr;
  hipGraphExec_t g_cycle = nullptr;   // one V-cycle (or IR sweep) + residual norm
  int g_cycle_mode = -1;
  // whole RK43 stages (state kernel, start, first V-cycle batch) by
  // (kind, stage, batch size), see mg_begin_stage
  std::map<int, hipGraphExec_t> g_stage;
  // V-cycles (and the learned guess's inference chain) are launched kernel by
  // kernel or as captured graphs.  Measured at 1025^2: one simulation steps
  // 3% faster with direct launches (a graph launch costs the GPU more than
  // the kernels' own launches, and the host keeps ahead of the GPU), while
  // an ensemble of 64 members driven from one host thread steps 12% faster
  // with graphs (there the host's launch work is the bound).  launch_env:
  // 0 = that choice, 1 = always direct (QDEQ_DIRECT_LAUNCH=1, for profilers
  // that cannot trace graph launches), 2 = always graphs
  // (QDEQ_GRAPH_LAUNCH=1); in_ensemble is set while qdeq_ensemble_step runs.
  int launch_env = 0;
  bool in_ensemble = false;
  bool host_start = false;     // start decisions on the host (see mg_begin)
  // Solves launched ahead (see spec_allowed in api.hip): the solve queued
  // behind the current one, whether it is still live (not abandoned by a
  // continuation of the current solve), the mailbox page of the kernels
  // being launched (two solves in flight write different pages), and
  // whether the launches in progress are such a solve's.
  qdeq::MgSolve mg_next;
  bool spec_pending = false;
  int mb_page = 0;
  bool spec_launch = false;
  // the gate word a solve launched ahead checks before it starts: 2 (the
  // solve before it converged) or 7 (k_combine accepted the attempt)
  int spec_go = 2;
  bool spec_off = false;       // QDEQ_SPECULATE=0
  long long spec_taken = 0, spec_abandoned = 0;
  long long spec_taken_next = 0;  // of them, stage 2 of the next step
  bool stage_graphs_off = false;  // QDEQ_STAGE_GRAPHS=0 (see stage_graphs)
  // Stop tests folded into the next V-cycle (see fold_fits): while a batch
  // is launched kernel by kernel, skip_gate drops the stop-test kernel after
  // a V-cycle that is not the batch's last, and fold_next makes the next
  // V-cycle's level-0 down kernel run that test
  bool skip_gate = false, fold_next = false;
  bool fold_off = false;        // QDEQ_FOLD_STOP=0
  long long folded_tests = 0;   // stop tests run in a next V-cycle
  int last_it = 2;             // V-cycles of the last solve (first batch size)
  // RK43 solves: the stage's key (stage s, +8 for the solve after an
  // accepted step) while its solve runs, else -1, and the V-cycles of the
  // last solve of each key (-1: none yet).  The V-cycles a stage needs
  // repeat from step to step far more closely than they repeat from one
  // stage to the next, so the first batch of a solve is sized by its key.
  int solve_key = -1;
  int last_it_key[16] = {-1, -1, -1, -1, -1, -1, -1, -1,
                         -1, -1, -1, -1, -1, -1, -1, -1};
  // the pending candidate start is in u's interior; the plain start's
  // interior is then at plain_interior (same layout as u)
  bool cand_in_u = false;
  const double *plain_interior = nullptr;
  // the next launched V-cycle carries the start of its solve (the start
  // norms and test folded into its down kernels, see mg_tiled.hip)
  bool fuse_start_next = false;
  std::string err;
  // RK43 host state (odeSolver.cpp:1338-1363)
  bool rk_ready = false;
  double t = 0, dt = 0, totErr = 0, errA[2] = {0, 0};
  int it = 0, attempts = 0;
  long long vcycles = 0, solves = 0, ir_sweeps = 0;
  double c_stress[4];  // partial1st coefficients (fault.cpp:162)
  std::vector<double> mu_fault;  // nodal mu at the fault column (host copy)
  // fast cycle: tiled kernels and their interior tile shape (even sizes)
  bool tiled = true;
  int tile_tx = 64, tile_ty = 32;
  // per-level shapes chosen by qdeq_autotune_tiles (tuned), else the above
  // with 16 thread rows
  bool tuned = false;
  qdeq::TileShape tdown[QDEQ_MAX_LEVELS] = {}, tup[QDEQ_MAX_LEVELS] = {};
  size_t lds_max = 65536;  // dynamic LDS per workgroup (queried at creation)
  // direct coarsest-level solve of the fast cycle (n_coarse == 0): inverse of
  // the coarsest operator, column-major, stored in FP32 (products and sums
  // in the solve's own precision)
  float *cinv32 = nullptr;
  int coarse_n = 0;
  // HIP-event timing of one kernel of the fast cycle (0 = off; see
  // qdeq_set_kernel_timing): on sampled V-cycles the cycle is launched
  // directly instead of as a graph (emit_events) with events around it
  int timing = 0;
  bool emit_events = false;
  // one pair per mailbox page: a solve launched ahead may hold a sample too
  hipEvent_t ev_t[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  hipEvent_t ev_caller = nullptr;  // the caller's null-stream work (QDEQ_DEVICE)
  double timed_ms = 0.0;
  long long timed_calls = 0, timed_seq = 0;
  // learned initial guess (guess.hip): attached network, mode, the response
  // W_R to unit remote data, inference scratch (16 padded rows) and the
  // boundary data of the previous solve
  qdeq_guess *guess = nullptr;
  int guess_mode = 0;
  double *wr = nullptr;
  float *gx = nullptr, *gz = nullptr, *gy = nullptr, *gsplit = nullptr;
  size_t gsplit_n = 0;
  __bf16 *gh[4] = {};
  double *gprev = nullptr, *gbefore = nullptr;
  double *gs = nullptr;  // [0] max|ds|, [1] remote change, [2] test factor
  bool gprev_valid = false, cand_pending = false, harvest_pending = false;
  // resumable solve and RK43 driver; mailbox polling state
  qdeq::MgSolve mg;
  qdeq::RkMachine rk;
  // batched V-cycles: device gate (stop, count, outcome, sequence), the gate
  // the kernels being launched check (set while capturing the V-cycle), the
  // sequence numbers issued, and the observed V-cycle contraction
  int *gate = nullptr;
  int *gate_ptr = nullptr;
  long long seq_issued = 0;
  double rho_est = 0.1;
  long long mb_polls = 0;
  // output records and event catalogue (output.hip)
  qdeq::OutputRing *out = nullptr;
  int *ev_state = nullptr;
  void *ev_table = nullptr;
  double *ev_slip0 = nullptr;
  double ev_thr = 0.0;
  int ev_cap = 0;
  bool ev_on = false;
  bool mb_idle = false;
  std::chrono::steady_clock::time_point mb_idle_since{};
  long long g_predicted = 0, g_accepted = 0, g_harvested = 0, g_seen = 0;
  hipGraphExec_t g_pred = nullptr;
  // projection start (proj.hip): ring of proj_slots stored solutions (N x P
  // each), their fault columns and remote values, the fit's coefficients,
  // and W_R
  int proj_slots = 0, proj_count = 0, proj_head = 0;
  double proj_inject = 1.0;  // test hook: scales the start's change
  double *proj_u = nullptr, *proj_g = nullptr, *proj_r = nullptr,
         *proj_coef = nullptr, *proj_wr = nullptr;
};

namespace qdeq {

int set_error(qdeq_ctx *c, int code, const std::string &msg);

// multigrid (mg.hip)
int mg_solve(qdeq_ctx *c, int *vcycles, double *res0, double *res);
int mg_begin(qdeq_ctx *c);
// stage graphs: whether this context replays whole RK stages, and the
// start of a stage solve as one graph replay (kind 0: stage s's state;
// kind 1: stage s - 1's fault update and stage s's state)
bool stage_graphs(const qdeq_ctx *c);
int mg_begin_stage(qdeq_ctx *c, int kind, int s);
int mg_begin_spec(qdeq_ctx *c, int go = 2);
int mg_poll(qdeq_ctx *c, bool block, int *done);
int mg_build_coefficients(qdeq_ctx *c);
void mg_destroy_graphs(qdeq_ctx *c);
bool use_tiled(const qdeq_ctx *c);
bool device_start(const qdeq_ctx *c);
bool device_start_cand(const qdeq_ctx *c);

void launch_resnorm_of(qdeq_ctx *c, const double *u, int off, int slot);
void mbox_arm(qdeq_ctx *c, int slot);
int mbox_wait(qdeq_ctx *c, int slot);

// output records and events (output.hip); output_after_step writes the
// periodic record when one is due
int output_after_step(qdeq_ctx *c);
void output_free(qdeq_ctx *c);
void events_after_step(qdeq_ctx *c);
void events_free(qdeq_ctx *c);

// learned initial guess (guess.hip)
int guess_before_solve(qdeq_ctx *c);
int guess_after_solve(qdeq_ctx *c);
void guess_reset(qdeq_ctx *c);
void guess_free_ctx(qdeq_ctx *c);
// W_R into dst: the solution for unit remote data and zero fault data
int solve_unit_remote(qdeq_ctx *c, double *dst);

// projection start (proj.hip)
int proj_before_solve(qdeq_ctx *c);
int proj_after_solve(qdeq_ctx *c);
void proj_reset(qdeq_ctx *c);
void proj_free(qdeq_ctx *c);

// tiled fast cycle (mg_tiled.hip)
int tail_entry_level(const qdeq_ctx *c);
bool tail_fits(const qdeq_ctx *c, bool fp64);
void vcycle_tiled_fp64(qdeq_ctx *c, double *partial, int *nparts);
void vcycle_tiled_fp32_ir(qdeq_ctx *c);
int tiled_setup(qdeq_ctx *c);
bool tiles_fit(const qdeq_ctx *c, int tx, int ty);
bool fused_start_fits(const qdeq_ctx *c);
bool fold_fits(const qdeq_ctx *c);
int autotune_tiles(qdeq_ctx *c, int reps);
int set_tile_shapes(qdeq_ctx *c, int nlevels, const int *shapes);
int kernel_breakdown(qdeq_ctx *c, int reps, int *n, double *us);

// marks the start (pos 0) or end (pos 1) of kernel `which` in the launch
// sequence: records the timing events on a directly launched V-cycle
inline void timing_mark(qdeq_ctx *c, int which, int pos) {
  if (c->emit_events && c->timing == which)
    (void)hipEventRecord(c->ev_t[c->mb_page][pos], c->stream);
}

// exact coarse solve (coarse.hip).  The matrix-vector product splits each
// row's columns into coarse_chunks(n) consecutive chunks of
// coarse_chunk_len(n) columns, summed in column order from zero, and adds
// the chunk sums in chunk order (one thread per row and chunk, <= 1024).
__host__ __device__ inline int coarse_chunks(int n) {
  const int q = 1024 / n;
  return q < 1 ? 1 : (q > 4 ? 4 : q);
}
__host__ __device__ inline int coarse_chunk_len(int n) {
  const int q = coarse_chunks(n);
  return (n + q - 1) / q;
}
// e[row] (rows < n, all threads of the block) = sum_col minvT[col*n+row] g[col]
// in the chunked order; part: >= 1024 LDS elements; ends with a barrier.
template <typename T>
__device__ inline void coarse_matvec(const float *__restrict__ minvT, const T *g,
                                     T *part, int n, T *out, int W, int K,
                                     size_t pitch) {
  const int Q = coarse_chunks(n), C = coarse_chunk_len(n);
  for (int t = threadIdx.x; t < Q * n; t += blockDim.x) {
    const int q = t / n, row = t - q * n;
    const int c0 = q * C, c1 = (q + 1) * C < n ? (q + 1) * C : n;
    T acc = T(0);
#pragma unroll 8
    for (int col = c0; col < c1; col++)
      acc = acc + (T)minvT[(size_t)col * n + row] * g[col];
    part[t] = acc;
  }
  __syncthreads();
  for (int row = threadIdx.x; row < n; row += blockDim.x) {
    T acc = part[row];
    for (int q = 1; q < Q; q++) acc = acc + part[q * n + row];
    const int j = row / W, k = row - j * W + 1;
    out[(size_t)j * pitch + k] = acc;
  }
  __syncthreads();
}

int build_coarse_inverse(qdeq_ctx *c);

// v of the lane that DPP control CTRL selects (both 32-bit halves moved).
template <int CTRL>
__device__ inline double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the 64 lanes of a wavefront (every lane active), returned to
// every lane: lane moves within rows of 16 (xor 1 and xor 2 in each quad,
// then the half-row and row mirrors; the two operands of each addition
// differ between lanes only in order, so every lane of a row holds the same
// bits), then the four row sums added in row order.
__device__ inline double wave_sum(double v) {
  v = v + dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]
  v = v + dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]
  v = v + dpp_move<0x141>(v);  // row_half_mirror
  v = v + dpp_move<0x140>(v);  // row_mirror
  const long long b = __double_as_longlong(v);
  double r[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int lo = __builtin_amdgcn_readlane((int)b, 16 * i);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 16 * i);
    r[i] = __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
  }
  return ((r[0] + r[1]) + r[2]) + r[3];
}

// Sums of NV values over the workgroup, returned to every thread: wavefront
// butterflies, then the wavefront sums added in wavefront order (one
// barrier).  sh: LDS for NV doubles per wavefront, not reused by the caller
// before a later barrier.
template <int NV>
__device__ inline void block_sums(double (&v)[NV], double *sh) {
  const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  const int nw = (blockDim.x * blockDim.y * blockDim.z + 63) >> 6;
#pragma unroll
  for (int q = 0; q < NV; q++) v[q] = wave_sum(v[q]);
  if ((tid & 63) == 0)
#pragma unroll
    for (int q = 0; q < NV; q++) sh[q * 16 + (tid >> 6)] = v[q];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; q++) {
    double a = sh[q * 16];
    for (int w = 1; w < nw; w++) a = a + sh[q * 16 + w];
    v[q] = a;
  }
}

__device__ inline double block_sum1(double v, double *sh) {
  double a[1] = {v};
  block_sums<1>(a, sh);
  return a[0];
}

// Entries of c->partial: the FP64 fast cycle's level-0 up kernel writes one
// norm partial per tile, so a level-0 tiling may have at most this many
// tiles (qdeq_set_tiles, set_tile_shapes and the autotuner enforce it).
constexpr int kPartialCap = 16384;

// Row-partial layout of the level-0 residual norms (mg.hip): one block per
// row, at most kNormCap blocks; plain norms at partial[0..), boundary data
// at partial[kNormCap..), a candidate start's norm at partial[2 kNormCap..).
constexpr int kNormCap = 2048;
inline int norm_blocks(const qdeq_ctx *c) {
  return c->lv[0].N < kNormCap ? c->lv[0].N : kNormCap;
}

// Host-mapped mailbox: two pages of SC_COUNT doubles; the kernels being
// launched, and the host reads of the solve being polled, use page mb_page.
inline double *mb_dev(const qdeq_ctx *c) { return c->mbox_dev + c->mb_page * SC_COUNT; }
inline double *mb_host(const qdeq_ctx *c) { return c->mbox_host + c->mb_page * SC_COUNT; }

inline bool launch_direct(const qdeq_ctx *c) {
  return c->launch_env == 1 || (c->launch_env == 0 && !c->in_ensemble);
}
void launch_coarse_direct_fp64(qdeq_ctx *c, double *u, const double *f);
void launch_coarse_direct_fp32(qdeq_ctx *c, float *u, const float *f);

// medium (medium.hip): from a host nodal array (allocates device staging),
// and from a device nodal array into a device output without any
// allocation (the weight tables in a file-scope device array)
int build_mu_equi_device(int ni, int nk, const double *mu_host, double *me_dev,
                         hipStream_t s, std::string *err);
int build_mu_equi_noalloc(int ni, int nk, const double *mu_dev, double *me_dev,
                          int device, std::string *err);

// fault + RK43 stages (fault.hip)
void launch_set_boundary_grid(qdeq_ctx *c, double *U, const double *slip,
                              const double *slip_shift, double t, double ratio,
                              double dt);
void launch_extend_crew(qdeq_ctx *c, double *U);
void launch_pack(qdeq_ctx *c, const double *U);
void launch_unpack(qdeq_ctx *c, double *U);
void launch_eval_stress_grid(qdeq_ctx *c, const double *U, const double *mu,
                             const double *tau_n0, double *Ts, double *Tn);
void launch_nr(qdeq_ctx *c, const double *Ts, const double *Tn,
               const double *psi, const double *a, const double *eta,
               double *V);
void launch_dpsi(qdeq_ctx *c, const double *psi, const double *V,
                 const double *b, double *dpsi);
// stage kernels: t and dt from scal[SC_T], scal[SC_DT]
void launch_stage(qdeq_ctx *c, int s, const int *go = nullptr);
// fault update of stage s and the stage state of s + 1 in one launch
void launch_update_stage(qdeq_ctx *c, int s, const int *spec = nullptr);
void launch_fault_update(qdeq_ctx *c, int s);
// zeroes the stage rates of a new attempt and sets its t, dt on the device
void launch_zero_stage_rates(qdeq_ctx *c, double t, double dt);
void launch_update_next(qdeq_ctx *c, double t, double dt, const int *spec);
void launch_combine(qdeq_ctx *c);
// accepts y4 and sets the device time to t_new
void launch_commit(qdeq_ctx *c, double t_new, const int *go = nullptr);
void launch_maxima(qdeq_ctx *c);

}  // namespace qdeq

#define QDEQ_HIP(c, call)                                                   \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      return qdeq::set_error((c), QDEQ_ERR_HIP,                             \
                             std::string(#call) + ": " +                    \
                                 hipGetErrorString(e_));                    \
  } while (0)
