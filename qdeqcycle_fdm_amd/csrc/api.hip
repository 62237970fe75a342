// api.hip -- context, drop-in entry points and the device-resident RK43
// driver behind include/qdeq.h.
//
// Host control flow follows XuSun-SUSTech/QDEqCycle_FDM RK43
// (odeSolver.cpp:1277-1674); the per-node and per-grid-point arithmetic runs
// in the kernels of fault.hip and mg.hip.  The reference's exit() paths
// become status codes.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "qdeq_internal.h"

namespace {
thread_local std::string g_error;
}

namespace qdeq {

int set_error(qdeq_ctx *c, int code, const std::string &msg) {
  if (c)
    c->err = msg;
  else
    g_error = msg;
  return code;
}

}  // namespace qdeq

using namespace qdeq;

static int begin_call(qdeq_ctx *c, int where);

template <typename T>
static hipError_t alloc_zero(T **p, size_t n, hipStream_t s) {
  hipError_t e = hipMalloc(p, n * sizeof(T));
  if (e != hipSuccess) return e;
  return hipMemsetAsync(*p, 0, n * sizeof(T), s);
}

extern "C" {

void qdeq_default_config(qdeq_config *c) {
  std::memset(c, 0, sizeof(*c));
  c->ni = 801;  // SetParams.cpp:21-23
  c->nk = 801;
  c->dh = 30.0;
  c->n_levels = 6;  // odeSolver.cpp:3
  c->n_pre = c->n_post = c->n_coarse = 32;  // odeSolver.cpp:4
  c->relax = 1.0;   // odeSolver.cpp:2
  c->cycle = QDEQ_CYCLE_REFERENCE;
  c->precision = QDEQ_PREC_FP64;
  c->tol_mode = QDEQ_TOL_ABS;
  c->tol = 1.0e-9;  // odeSolver.cpp:689
  c->max_vcycles = 1000;
  c->ini_dt = 1e-3;  // SetParams.cpp:14-17
  c->mindt = 1e-5;
  c->maxdt = 1e8;
  c->totTol = 1e-7;
  c->tmax = 0 + 60.0 * (1 + 60.0 * (0 + 24.0 * (0 + 30.0 * (0 + 12.0 * 3000))));
  c->f0 = 0.6;  // fault.cpp:85-89
  c->V0 = 1e-6;
  c->L = 0.008;
  c->Vini = 1e-9;
  c->device = 0;
}

const char *qdeq_last_error(const qdeq_ctx *ctx) {
  return ctx ? ctx->err.c_str() : g_error.c_str();
}

const char *qdeq_global_error(void) { return g_error.c_str(); }

int qdeq_hip_runtime_version(int *built, int *running) {
  if (built) *built = HIP_VERSION;
  if (!running) return QDEQ_OK;
  *running = 0;
  const hipError_t e = hipRuntimeGetVersion(running);
  if (e != hipSuccess) return set_error(nullptr, QDEQ_ERR_HIP, hipGetErrorString(e));
  return QDEQ_OK;
}

static void free_ctx(qdeq_ctx *c) {
  if (!c) return;
  mg_destroy_graphs(c);
  for (int l = 0; l < QDEQ_MAX_LEVELS; l++) {
    (void)hipFree(c->d[l].u); (void)hipFree(c->d[l].f); (void)hipFree(c->d[l].r);
    (void)hipFree(c->d[l].cx); (void)hipFree(c->d[l].cz);
    (void)hipFree(c->s[l].u); (void)hipFree(c->s[l].f); (void)hipFree(c->s[l].r);
    (void)hipFree(c->s[l].cx); (void)hipFree(c->s[l].cz);
  }
  (void)hipFree(c->mu_equi);
  (void)hipFree(c->cinv32);
  (void)hipFree(c->partial);
  (void)hipFree(c->gate);
  (void)hipFree(c->scal);
  (void)hipFree(c->flags);
  (void)hipFree(c->stage_U);
  for (auto &p : c->stage_v) (void)hipFree(p);
  (void)hipFree(c->fb.base);
  if (c->scal_host) (void)hipHostFree(c->scal_host);
  if (c->mbox_host) (void)hipHostFree(c->mbox_host);
  if (c->flags_host) (void)hipHostFree(c->flags_host);
  for (auto &pair : c->ev_t)
    for (hipEvent_t e : pair)
      if (e) (void)hipEventDestroy(e);
  guess_free_ctx(c);
  proj_free(c);
  (void)qdeq_output_close(c);
  events_free(c);
  if (c->rk.ev) (void)hipEventDestroy(c->rk.ev);
  if (c->ev_caller) (void)hipEventDestroy(c->ev_caller);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

#define CREATE_HIP(call)                                                   \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess) {                                                \
      set_error(nullptr, QDEQ_ERR_HIP,                                     \
                std::string(#call) + ": " + hipGetErrorString(e_));        \
      free_ctx(c);                                                         \
      return QDEQ_ERR_HIP;                                                 \
    }                                                                      \
  } while (0)

int qdeq_create(const qdeq_config *cfg, const double *mu,
                const double *mu_equi, qdeq_ctx **out) {
  if (!cfg || !mu || !out)
    return set_error(nullptr, QDEQ_ERR_ARG, "qdeq_create: null argument");
  *out = nullptr;
  if (cfg->ni < 5 || cfg->nk < 5 || cfg->dh <= 0 || cfg->n_levels < 1 ||
      cfg->n_levels > QDEQ_MAX_LEVELS || cfg->n_pre < 0 || cfg->n_post < 0 ||
      cfg->n_coarse < 0 ||
      (cfg->n_coarse == 0 && cfg->cycle != QDEQ_CYCLE_FAST) ||
      cfg->max_vcycles < 1 || cfg->tol <= 0)
    return set_error(nullptr, QDEQ_ERR_ARG, "qdeq_create: invalid config");
  qdeq_ctx *c = new qdeq_ctx();
  c->cfg = *cfg;
  c->nx = cfg->ni + 2;
  c->nz = cfg->nk + 2;
  c->L = cfg->n_levels;
  c->tile_tx = 64;
  c->tile_ty = cfg->precision == QDEQ_PREC_FP32_IR ? 32 : 16;
  {
    int n = cfg->nk, k = cfg->ni, fc = 1;
    double h = cfg->dh;
    for (int l = 0; l < c->L; l++) {
      if (n < 3 || k < 3) {
        delete c;
        return set_error(nullptr, QDEQ_ERR_ARG,
                         "qdeq_create: too many levels for the grid");
      }
      c->lv[l] = LevelDims{n, k, (k + 63) / 64 * 64, fc, h};
      n = (n - 1) / 2 + 1;
      k = (k - 1) / 2 + 1;
      h *= 2;
      fc *= 2;
    }
  }
  const double dh = cfg->dh;  // partial1st (fault.cpp:162)
  c->c_stress[0] = -24.0 / (17.0 * dh);
  c->c_stress[1] = 59.0 / (34.0 * dh);
  c->c_stress[2] = -4.0 / (17.0 * dh);
  c->c_stress[3] = -3.0 / (34.0 * dh);

  CREATE_HIP(hipSetDevice(cfg->device));
  CREATE_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (int l = 0; l < c->L; l++) {
    const LevelDims &d = c->lv[l];
    const size_t n = (size_t)d.N * d.P, nz = (size_t)(d.N + 1) * d.P;
    CREATE_HIP(alloc_zero(&c->d[l].u, n, c->stream));
    CREATE_HIP(alloc_zero(&c->d[l].f, n, c->stream));
    CREATE_HIP(alloc_zero(&c->d[l].r, n, c->stream));
    CREATE_HIP(alloc_zero(&c->d[l].cx, n, c->stream));
    CREATE_HIP(alloc_zero(&c->d[l].cz, nz, c->stream));
    CREATE_HIP(alloc_zero(&c->s[l].u, n, c->stream));
    CREATE_HIP(alloc_zero(&c->s[l].f, n, c->stream));
    CREATE_HIP(alloc_zero(&c->s[l].r, n, c->stream));
    CREATE_HIP(alloc_zero(&c->s[l].cx, n, c->stream));
    CREATE_HIP(alloc_zero(&c->s[l].cz, nz, c->stream));
  }
  const size_t nme = (size_t)(2 * c->nx - 1) * (2 * c->nz - 1);
  CREATE_HIP(hipMalloc(&c->mu_equi, nme * sizeof(double)));
  if (mu_equi) {
    CREATE_HIP(hipMemcpyAsync(c->mu_equi, mu_equi, nme * sizeof(double),
                              hipMemcpyHostToDevice, c->stream));
  } else {
    std::string e;
    if (build_mu_equi_device(cfg->ni, cfg->nk, mu, c->mu_equi, c->stream, &e) !=
        QDEQ_OK) {
      set_error(nullptr, QDEQ_ERR_HIP, "mu_equi: " + e);
      free_ctx(c);
      return QDEQ_ERR_HIP;
    }
  }
  CREATE_HIP(alloc_zero(&c->partial, kPartialCap, c->stream));
  CREATE_HIP(alloc_zero(&c->scal, SC_COUNT, c->stream));
  CREATE_HIP(alloc_zero(&c->flags, 4, c->stream));
  CREATE_HIP(hipHostMalloc(&c->scal_host, SC_COUNT * sizeof(double)));
  CREATE_HIP(hipHostMalloc(&c->mbox_host, 2 * SC_COUNT * sizeof(double),
                           hipHostMallocMapped | hipHostMallocCoherent));
  CREATE_HIP(hipHostGetDevicePointer((void **)&c->mbox_dev, c->mbox_host, 0));
  std::memset(c->mbox_host, 0, 2 * SC_COUNT * sizeof(double));
  CREATE_HIP(alloc_zero(&c->gate, 8, c->stream));
  CREATE_HIP(hipHostMalloc(&c->flags_host, 4 * sizeof(int)));
  CREATE_HIP(alloc_zero(&c->stage_U, (size_t)2 * c->nx * c->nz, c->stream));
  for (auto &p : c->stage_v) CREATE_HIP(alloc_zero(&p, (size_t)c->nz, c->stream));

  // fault / RK43 buffers: 40 vectors of nz, then 14 device pointers
  const size_t nzs = (size_t)c->nz;
  CREATE_HIP(hipMalloc(&c->fb.base, (40 * nzs + 14) * sizeof(double)));
  CREATE_HIP(hipMemsetAsync(c->fb.base, 0, 40 * nzs * sizeof(double), c->stream));
  {
    FaultBufs &f = c->fb;
    double *b = f.base;
    double **slots[] = {&f.slip, &f.psi, &f.slip_shift, &f.a, &f.b, &f.eta,
                        &f.tau_n0, &f.mu_f, &f.Ts, &f.Tn};
    int q = 0;
    for (double **s : slots) *s = b + (q++) * nzs;
    for (int s = 1; s <= 6; s++) f.Ys[s] = b + (q++) * nzs;
    for (int s = 1; s <= 6; s++) f.Yp[s] = b + (q++) * nzs;
    for (int s = 1; s <= 6; s++) f.FV[s] = b + (q++) * nzs;
    for (int s = 1; s <= 6; s++) f.Fd[s] = b + (q++) * nzs;
    f.y3s = b + (q++) * nzs;
    f.y3p = b + (q++) * nzs;
    f.y4s = b + (q++) * nzs;
    f.y4p = b + (q++) * nzs;
    f.Ys[0] = f.Yp[0] = f.FV[0] = f.Fd[0] = nullptr;
    double *ptrs[14];
    for (int s = 0; s < 7; s++) {
      ptrs[s] = f.FV[s];
      ptrs[7 + s] = f.Fd[s];
    }
    CREATE_HIP(hipMemcpyAsync(b + 40 * nzs, ptrs, sizeof(ptrs),
                              hipMemcpyHostToDevice, c->stream));
    c->mu_fault.assign(nzs, 0.0);
    for (int k = 1; k <= cfg->nk; k++) c->mu_fault[k] = mu[(size_t)k * c->nx + 1];
    CREATE_HIP(hipMemcpyAsync(f.mu_f, c->mu_fault.data(), nzs * sizeof(double),
                              hipMemcpyHostToDevice, c->stream));
    CREATE_HIP(hipStreamSynchronize(c->stream));
  }
  if (mg_build_coefficients(c) != QDEQ_OK || build_coarse_inverse(c) != QDEQ_OK ||
      tiled_setup(c) != QDEQ_OK) {
    set_error(nullptr, QDEQ_ERR_HIP, c->err);
    free_ctx(c);
    return QDEQ_ERR_HIP;
  }
  // QDEQ_DIRECT_LAUNCH=1 / QDEQ_GRAPH_LAUNCH=1: V-cycles always launched
  // kernel by kernel / always as the captured graph (same kernels and
  // arguments); by default direct, graphs inside qdeq_ensemble_step
  const char *dl = std::getenv("QDEQ_DIRECT_LAUNCH");
  const char *gl = std::getenv("QDEQ_GRAPH_LAUNCH");
  c->launch_env = (dl && dl[0] == '1') ? 1 : ((gl && gl[0] == '1') ? 2 : 0);
  // QDEQ_HOST_START=1: the host reads every solve's start norms and decides
  // (the path FP32-IR solves always take) instead of the device
  const char *hs = std::getenv("QDEQ_HOST_START");
  c->host_start = hs && hs[0] == '1';
  // QDEQ_STAGE_GRAPHS=0: graph launches replay single V-cycles only (the
  // stage's state, start and first batch are launched one by one)
  const char *sg = std::getenv("QDEQ_STAGE_GRAPHS");
  c->stage_graphs_off = sg && sg[0] == '0';
  // QDEQ_SPECULATE=0: every stage waits for the previous solve's end
  const char *sp = std::getenv("QDEQ_SPECULATE");
  c->spec_off = sp && sp[0] == '0';
  // QDEQ_FOLD_STOP=0: every V-cycle of a batch has its own stop-test kernel
  const char *fs = std::getenv("QDEQ_FOLD_STOP");
  c->fold_off = fs && fs[0] == '0';
  *out = c;
  return QDEQ_OK;
}

int qdeq_destroy(qdeq_ctx *ctx) {
  free_ctx(ctx);
  return QDEQ_OK;
}

int qdeq_set_tiles(qdeq_ctx *c, int tiled, int tile_x, int tile_y) {
  if (!c) return set_error(c, QDEQ_ERR_ARG, "null argument");
  if (tile_x < 8 || tile_y < 2 || (tile_x & 1) || (tile_y & 1) || tile_x > 256 ||
      tile_y > 128)
    return set_error(c, QDEQ_ERR_ARG, "tile sizes must be even, 8..256 x 2..128");
  if (!tiles_fit(c, tile_x, tile_y))
    return set_error(c, QDEQ_ERR_ARG, "tile too large for the LDS at these sweep counts");
  if (c->cfg.precision == QDEQ_PREC_FP64) {
    // the FP64 cycle's level-0 up kernel writes one norm partial per tile
    const LevelDims &d = c->lv[0];
    const long blocks = (long)((d.K + tile_x - 1) / tile_x) * ((d.N + tile_y - 1) / tile_y);
    if (blocks > kPartialCap)
      return set_error(c, QDEQ_ERR_ARG,
                       "tile too small: " + std::to_string(blocks) +
                           " level-0 tiles, at most " + std::to_string(kPartialCap));
  }
  c->tiled = tiled != 0;
  c->tile_tx = tile_x;
  c->tile_ty = tile_y;
  c->tuned = false;
  mg_destroy_graphs(c);
  return QDEQ_OK;
}

int qdeq_autotune_tiles(qdeq_ctx *c, int reps) {
  if (!c || reps < 1) return set_error(c, QDEQ_ERR_ARG, "bad argument");
  return autotune_tiles(c, reps);
}

int qdeq_set_tile_shapes(qdeq_ctx *c, int nlevels, const int *shapes) {
  if (!c || !shapes) return set_error(c, QDEQ_ERR_ARG, "null argument");
  return set_tile_shapes(c, nlevels, shapes);
}

int qdeq_kernel_breakdown(qdeq_ctx *c, int reps, int *n, double *us) {
  if (!c || reps < 1 || !n || !us) return set_error(c, QDEQ_ERR_ARG, "bad argument");
  return kernel_breakdown(c, reps, n, us);
}

int qdeq_get_tile_shapes(const qdeq_ctx *c, int *nlevels, int *shapes) {
  if (!c || !nlevels || !shapes) return QDEQ_ERR_ARG;
  const int lt = use_tiled(c) ? tail_entry_level(c) : 0;
  *nlevels = lt;
  for (int l = 0; l < lt; l++) {
    const TileShape dn = c->tuned ? c->tdown[l] : TileShape{c->tile_tx, c->tile_ty, 16};
    const TileShape up = c->tuned ? c->tup[l] : TileShape{c->tile_tx, c->tile_ty, 16};
    const int v[6] = {dn.tx, dn.ty, dn.tyt, up.tx, up.ty, up.tyt};
    for (int i = 0; i < 6; i++) shapes[6 * l + i] = v[i];
  }
  return QDEQ_OK;
}

int qdeq_uses_tiled_cycle(const qdeq_ctx *c) { return c && use_tiled(c) ? 1 : 0; }

int qdeq_device_synchronize(int device) {
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) return set_error(nullptr, QDEQ_ERR_HIP, hipGetErrorString(e));
  return QDEQ_OK;
}

int qdeq_set_kernel_timing(qdeq_ctx *c, int which) {
  if (!c || which < 0 || which > QDEQ_TIME_FAULT)
    return set_error(c, QDEQ_ERR_ARG, "bad timing selector");
  if (which && !c->ev_t[0][0]) {
    for (auto &pair : c->ev_t)
      for (hipEvent_t &e : pair) QDEQ_HIP(c, hipEventCreate(&e));
  }
  c->timing = which;
  c->timed_ms = 0.0;
  c->timed_calls = 0;
  c->timed_seq = 0;
  mg_destroy_graphs(c);
  return QDEQ_OK;
}

int qdeq_get_kernel_timing(qdeq_ctx *c, double *total_ms, long long *calls) {
  if (!c) return set_error(c, QDEQ_ERR_ARG, "null argument");
  if (total_ms) *total_ms = c->timed_ms;
  if (calls) *calls = c->timed_calls;
  return QDEQ_OK;
}

__global__ void k_copy_probe(const double *__restrict__ a, double *__restrict__ b,
                             size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = a[i];
}

int qdeq_copy_probe(qdeq_ctx *c, long long n_doubles, int reps, double *avg_ms) {
  if (!c || n_doubles <= 0 || reps <= 0)
    return set_error(c, QDEQ_ERR_ARG, "bad copy probe size");
  double *a = nullptr, *b = nullptr;
  hipEvent_t e0, e1;
  QDEQ_HIP(c, hipMalloc(&a, (size_t)n_doubles * sizeof(double)));
  QDEQ_HIP(c, hipMalloc(&b, (size_t)n_doubles * sizeof(double)));
  QDEQ_HIP(c, hipMemsetAsync(a, 0, (size_t)n_doubles * sizeof(double), c->stream));
  QDEQ_HIP(c, hipEventCreate(&e0));
  QDEQ_HIP(c, hipEventCreate(&e1));
  const unsigned blocks = 256 * 8;
  hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, c->stream, a, b,
                     (size_t)n_doubles);
  QDEQ_HIP(c, hipEventRecord(e0, c->stream));
  for (int r = 0; r < reps; r++)
    hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, c->stream, a, b,
                       (size_t)n_doubles);
  QDEQ_HIP(c, hipEventRecord(e1, c->stream));
  QDEQ_HIP(c, hipEventSynchronize(e1));
  float ms = 0.0f;
  QDEQ_HIP(c, hipEventElapsedTime(&ms, e0, e1));
  if (avg_ms) *avg_ms = ms / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return QDEQ_OK;
}

int qdeq_get_mu_equi(qdeq_ctx *c, double *me, int where) {
  if (!c || !me) return set_error(c, QDEQ_ERR_ARG, "null argument");
  if (int st = begin_call(c, where)) return st;
  const size_t nme = (size_t)(2 * c->nx - 1) * (2 * c->nz - 1);
  QDEQ_HIP(c, hipMemcpyAsync(me, c->mu_equi, nme * sizeof(double),
                             where == QDEQ_HOST ? hipMemcpyDeviceToHost
                                                : hipMemcpyDeviceToDevice,
                             c->stream));
  QDEQ_HIP(c, hipStreamSynchronize(c->stream));
  return QDEQ_OK;
}

int qdeq_build_mu_equi(int ni, int nk, const double *mu, double *me,
                       int device) {
  if (ni < 3 || nk < 3 || !mu || !me)
    return set_error(nullptr, QDEQ_ERR_ARG, "qdeq_build_mu_equi: bad argument");
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess)
    return set_error(nullptr, QDEQ_ERR_HIP, hipGetErrorString(e));
  const size_t nme = (size_t)(2 * (ni + 2) - 1) * (2 * (nk + 2) - 1);
  double *d = nullptr;
  hipStream_t s = nullptr;
  e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&d, nme * sizeof(double));
  if (e != hipSuccess) {
    if (s) (void)hipStreamDestroy(s);
    return set_error(nullptr, QDEQ_ERR_HIP, hipGetErrorString(e));
  }
  std::string msg;
  int st = build_mu_equi_device(ni, nk, mu, d, s, &msg);
  if (st == QDEQ_OK) {
    e = hipMemcpyAsync(me, d, nme * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      msg = hipGetErrorString(e);
      st = QDEQ_ERR_HIP;
    }
  }
  (void)hipFree(d);
  (void)hipStreamDestroy(s);
  if (st != QDEQ_OK) return set_error(nullptr, st, msg);
  return QDEQ_OK;
}

// Largest grid side of the allocation-free mu_equi build: its face kernels
// launch one grid row per node row, and a launch has at most 65535 rows.
constexpr int kNoallocMaxSide = 65535;

int qdeq_build_mu_equi_noalloc(int ni, int nk, const double *mu, double *me,
                               int device) {
  if (!mu || !me)
    return set_error(nullptr, QDEQ_ERR_ARG, "qdeq_build_mu_equi_noalloc: null argument");
  if (ni < 3 || nk < 3 || ni > kNoallocMaxSide || nk > kNoallocMaxSide)
    return set_error(nullptr, QDEQ_ERR_ARG,
                     "qdeq_build_mu_equi_noalloc: grid " + std::to_string(ni) + " x " +
                         std::to_string(nk) + " outside 3.." +
                         std::to_string(kNoallocMaxSide) + " nodes per side");
  std::string msg;
  const int st = build_mu_equi_noalloc(ni, nk, mu, me, device, &msg);
  if (st != QDEQ_OK)
    return set_error(nullptr, st, "qdeq_build_mu_equi_noalloc: " + msg);
  return QDEQ_OK;
}

// ---- drop-in entry points ----

// Arrays passed as QDEQ_DEVICE may have been written by the caller's work on
// the null stream; the context's own stream does not wait for that stream by
// itself (it is non-blocking), so it waits for an event recorded there.
static int begin_call(qdeq_ctx *c, int where) {
  if (where != QDEQ_HOST && where != QDEQ_DEVICE)
    return set_error(c, QDEQ_ERR_ARG, "pointer space must be QDEQ_HOST or QDEQ_DEVICE");
  if (where == QDEQ_HOST) return QDEQ_OK;
  if (!c->ev_caller)
    QDEQ_HIP(c, hipEventCreateWithFlags(&c->ev_caller, hipEventDisableTiming));
  QDEQ_HIP(c, hipEventRecord(c->ev_caller, nullptr));
  QDEQ_HIP(c, hipStreamWaitEvent(c->stream, c->ev_caller, 0));
  return QDEQ_OK;
}

static int to_device(qdeq_ctx *c, const double *src, double *stage, size_t n,
                     int where, const double **dst) {
  if (where == QDEQ_DEVICE) {
    *dst = src;
    return QDEQ_OK;
  }
  QDEQ_HIP(c, hipMemcpyAsync(stage, src, n * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
  *dst = stage;
  return QDEQ_OK;
}

static int to_host(qdeq_ctx *c, double *dst, const double *stage, size_t n,
                   int where) {
  if (where == QDEQ_DEVICE) return QDEQ_OK;
  QDEQ_HIP(c, hipMemcpyAsync(dst, stage, n * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
  return QDEQ_OK;
}

int qdeq_ode_solve(qdeq_ctx *c, double *U, int where, int *vcycles,
                   double *res0, double *res) {
  if (!c || !U) return set_error(c, QDEQ_ERR_ARG, "null argument");
  const size_t ng = (size_t)c->nx * c->nz;
  const double *Ud;
  int st = begin_call(c, where);
  if (st) return st;
  st = to_device(c, U, c->stage_U, ng, where, &Ud);
  if (st) return st;
  double *Uw = const_cast<double *>(Ud);
  launch_pack(c, Uw);
  st = mg_solve(c, vcycles, res0, res);
  launch_unpack(c, Uw);
  launch_extend_crew(c, Uw);
  QDEQ_HIP(c, hipGetLastError());
  int st2 = to_host(c, U, c->stage_U, ng, where);
  QDEQ_HIP(c, hipStreamSynchronize(c->stream));
  return st ? st : st2;
}

int qdeq_set_boundary(qdeq_ctx *c, double *U, const double *slip,
                      const double *ss, double t, double ratio, double dt,
                      int where) {
  if (!c || !U || !slip || !ss) return set_error(c, QDEQ_ERR_ARG, "null argument");
  const size_t ng = (size_t)c->nx * c->nz, nz = (size_t)c->nz;
  const double *Ud, *sd, *ssd;
  int st;
  if ((st = begin_call(c, where))) return st;
  if ((st = to_device(c, U, c->stage_U, ng, where, &Ud))) return st;
  if ((st = to_device(c, slip, c->stage_v[0], nz, where, &sd))) return st;
  if ((st = to_device(c, ss, c->stage_v[1], nz, where, &ssd))) return st;
  launch_set_boundary_grid(c, const_cast<double *>(Ud), sd, ssd, t, ratio, dt);
  QDEQ_HIP(c, hipGetLastError());
  if ((st = to_host(c, U, c->stage_U, ng, where))) return st;
  QDEQ_HIP(c, hipStreamSynchronize(c->stream));
  return QDEQ_OK;
}

int qdeq_eval_stress(qdeq_ctx *c, const double *U, const double *mu,
                     const double *tau_n0, double *Ts, double *Tn, int where) {
  if (!c || !U || !mu || !tau_n0 || !Ts || !Tn)
    return set_error(c, QDEQ_ERR_ARG, "null argument");
  const size_t ng = (size_t)c->nx * c->nz, nz = (size_t)c->nz;
  const double *Ud, *mud, *td;
  int st;
  if ((st = begin_call(c, where))) return st;
  if ((st = to_device(c, U, c->stage_U, ng, where, &Ud))) return st;
  if ((st = to_device(c, mu, c->stage_U + ng, ng, where, &mud))) return st;
  if ((st = to_device(c, tau_n0, c->stage_v[0], nz, where, &td))) return st;
  double *Tsd = where == QDEQ_DEVICE ? Ts : c->stage_v[1];
  double *Tnd = where == QDEQ_DEVICE ? Tn : c->stage_v[2];
  if (where != QDEQ_DEVICE) {
    QDEQ_HIP(c, hipMemcpyAsync(Tsd, Ts, nz * sizeof(double),
                               hipMemcpyHostToDevice, c->stream));
    QDEQ_HIP(c, hipMemcpyAsync(Tnd, Tn, nz * sizeof(double),
                               hipMemcpyHostToDevice, c->stream));
  }
  launch_eval_stress_grid(c, Ud, mud, td, Tsd, Tnd);
  QDEQ_HIP(c, hipGetLastError());
  if ((st = to_host(c, Ts, Tsd, nz, where))) return st;
  if ((st = to_host(c, Tn, Tnd, nz, where))) return st;
  QDEQ_HIP(c, hipStreamSynchronize(c->stream));
  return QDEQ_OK;
}

int qdeq_nr_search(qdeq_ctx *c, const double *Ts, const double *Tn,
                   const double *psi, const double *a, const double *eta,
                   double *V, int where) {
  if (!c || !Ts || !Tn || !psi || !a || !eta || !V)
    return set_error(c, QDEQ_ERR_ARG, "null argument");
  const size_t nz = (size_t)c->nz;
  const double *p[5];
  const double *src[5] = {Ts, Tn, psi, a, eta};
  int st;
  if ((st = begin_call(c, where))) return st;
  for (int i = 0; i < 5; i++)
    if ((st = to_device(c, src[i], c->stage_v[i], nz, where, &p[i]))) return st;
  const double *Vd0;
  if ((st = to_device(c, V, c->stage_v[5], nz, where, &Vd0))) return st;
  double *Vd = const_cast<double *>(Vd0);
  QDEQ_HIP(c, hipMemsetAsync(c->flags, 0, sizeof(int), c->stream));
  launch_nr(c, p[0], p[1], p[2], p[3], p[4], Vd);
  QDEQ_HIP(c, hipGetLastError());
  if ((st = to_host(c, V, Vd, nz, where))) return st;
  QDEQ_HIP(c, hipMemcpyAsync(c->flags_host, c->flags, sizeof(int),
                             hipMemcpyDeviceToHost, c->stream));
  QDEQ_HIP(c, hipStreamSynchronize(c->stream));
  if (c->flags_host[0])
    return set_error(c, QDEQ_ERR_NAN_FRICTION, "NaN, error for friction");
  return QDEQ_OK;
}

int qdeq_cal_dpsi(qdeq_ctx *c, const double *psi, const double *V,
                  const double *b, double *dpsi, int where) {
  if (!c || !psi || !V || !b || !dpsi)
    return set_error(c, QDEQ_ERR_ARG, "null argument");
  const size_t nz = (size_t)c->nz;
  const double *pd, *vd, *bd;
  int st;
  if ((st = begin_call(c, where))) return st;
  if ((st = to_device(c, psi, c->stage_v[0], nz, where, &pd))) return st;
  if ((st = to_device(c, V, c->stage_v[1], nz, where, &vd))) return st;
  if ((st = to_device(c, b, c->stage_v[2], nz, where, &bd))) return st;
  double *dd = where == QDEQ_DEVICE ? dpsi : c->stage_v[3];
  if (where != QDEQ_DEVICE)
    QDEQ_HIP(c, hipMemcpyAsync(dd, dpsi, nz * sizeof(double),
                               hipMemcpyHostToDevice, c->stream));
  launch_dpsi(c, pd, vd, bd, dd);
  QDEQ_HIP(c, hipGetLastError());
  if ((st = to_host(c, dpsi, dd, nz, where))) return st;
  QDEQ_HIP(c, hipStreamSynchronize(c->stream));
  return QDEQ_OK;
}

double qdeq_compute_error(int nk, const double *slip3, const double *psi3,
                          const double *slip4, const double *psi4) {
  double err1 = 0.0, err2 = 0.0;
  for (int k = 1; k <= nk; k++) {
    const double t1 = slip3[k] - slip4[k];
    const double t2 = psi3[k] - psi4[k];
    err1 = err1 + t1 * t1;
    err2 = err2 + t2 * t2;
  }
  err1 = std::sqrt(err1);
  err2 = std::sqrt(err2);
  return err1 / (std::sqrt((double)nk) * 1.0) + err2 / (std::sqrt((double)nk) * 1.0);
}

double qdeq_compute_step_size(int it, double dt, double mindt, double maxdt,
                              double totErr, double totTol,
                              const double *errA) {
  const double ord = 4.0, kappa = 0.9;
  const double alpha = 0.49 / ord, beta = 0.34 / ord, gamma = 0.1 / ord;
  double r;
  if (it < 4)
    r = kappa * std::pow(totTol / totErr, 1.0 / (1.0 + ord));
  else
    r = kappa * std::pow(totTol / totErr, alpha) * std::pow(errA[0] / totTol, beta) *
        std::pow(totTol / errA[1], gamma);
  double dT = r * dt;
  dT = (dt * 5.0 <= dT) ? dt * 5.0 : dT;
  dT = (maxdt <= dT) ? maxdt : dT;
  dT = (mindt >= dT) ? mindt : dT;
  return dT;
}

// ---- RK43 driver ----

static int check_flags(qdeq_ctx *c) {
  if (c->flags_host[0])
    return set_error(c, QDEQ_ERR_NAN_FRICTION, "NaN, error for friction");
  return QDEQ_OK;
}

int qdeq_rk43_setup(qdeq_ctx *c, const double *a, const double *b,
                    const double *eta, const double *tau_n0,
                    const double *U0) {
  if (!c || !a || !b || !eta || !tau_n0 || !U0)
    return set_error(c, QDEQ_ERR_ARG, "null argument");
  const int nk = c->cfg.nk, ni = c->cfg.ni;
  const size_t nz = (size_t)c->nz, nx = (size_t)c->nx;
  FaultBufs &f = c->fb;
  // fault_init (fault.cpp:116-133)
  std::vector<double> slip(nz, 0.0), psi(nz, 0.0), ss(nz, 0.0);
  for (int k = 1; k <= nk; k++) {
    psi[k] = std::log(c->cfg.V0 / c->cfg.Vini);
    psi[k] = c->cfg.f0 + b[k] * psi[k];
  }
  for (int k = 1; k <= nk; k++) slip[k] = 2.0 * U0[k * nx + 1];
  for (size_t k = 0; k < nz; k++) ss[k] = U0[k * nx + ni];
  QDEQ_HIP(c, hipMemsetAsync(f.base, 0, 40 * nz * sizeof(double), c->stream));
  struct {
    double *dst;
    const double *src;
  } up[] = {{f.a, a},       {f.b, b},         {f.eta, eta}, {f.tau_n0, tau_n0},
            {f.slip, slip.data()}, {f.psi, psi.data()}, {f.slip_shift, ss.data()}};
  for (auto &u : up)
    QDEQ_HIP(c, hipMemcpyAsync(u.dst, u.src, nz * sizeof(double),
                               hipMemcpyHostToDevice, c->stream));
  QDEQ_HIP(c, hipMemcpyAsync(f.mu_f, c->mu_fault.data(), nz * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
  QDEQ_HIP(c, hipMemcpyAsync(c->stage_U, U0, nx * nz * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
  launch_pack(c, c->stage_U);
  // the t = 0 stage reads t and dt on the device
  const std::vector<double> tdt{0.0, c->cfg.ini_dt};
  QDEQ_HIP(c, hipMemcpyAsync(c->scal + SC_T, tdt.data(), 2 * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
  QDEQ_HIP(c, hipMemsetAsync(c->flags, 0, sizeof(int), c->stream));
  QDEQ_HIP(c, hipGetLastError());
  // the copies read host vectors local to this call (and the caller's
  // arrays): they must have completed before it returns
  QDEQ_HIP(c, hipStreamSynchronize(c->stream));
  c->t = 0.0;
  c->dt = c->cfg.ini_dt;  // odeSolver.cpp:1363
  c->it = 0;
  c->attempts = 0;
  c->errA[0] = c->errA[1] = 0.0;
  c->totErr = 0.0;
  c->vcycles = c->solves = c->ir_sweeps = 0;
  c->spec_taken = c->spec_abandoned = c->spec_taken_next = 0;
  c->rk_ready = false;
  c->rk.phase = 0;  // idle
  c->rk.staged = false;
  c->rk.spec = false;
  c->spec_pending = false;
  c->mb_page = 0;
  c->mg.phase = 0;
  guess_reset(c);
  proj_reset(c);
  return QDEQ_OK;
}

// ---- resumable RK43 driver (odeSolver.cpp:1428-1633) ----
//
// The step loop of the reference as a state machine: every wait (a solve's
// norms, the embedded error and the NaN flag after the combination) is a
// phase boundary, so one host thread can advance several contexts, each on
// its own stream, without blocking on any of them (qdeq_ensemble_step).
// qdeq_rk43_step drives the same machine with blocking waits, so a context
// launches the same work in the same order either way.
enum {
  RK_IDLE = 0,
  RK_FIRST,          // stage 1 at t = 0 from y_n (odeSolver.cpp:1428-1437)
  RK_STEP_BEGIN,
  RK_ATTEMPT_BEGIN,
  RK_STAGE_BEGIN,
  RK_SOLVING,
  RK_WAIT_COMBINE,
  RK_WAIT_FIRST,
  RK_STEP_FINISH,
  RK_WAIT_END,
  RK_DONE
};

// Host readback of n ints of flags (and the embedded error) + an event
static int rk_readback(qdeq_ctx *c, bool toterr) {
  RkMachine &r = c->rk;
  if (!r.ev) QDEQ_HIP(c, hipEventCreateWithFlags(&r.ev, hipEventDisableTiming));
  QDEQ_HIP(c, hipGetLastError());
  if (toterr)
    QDEQ_HIP(c, hipMemcpyAsync(c->scal_host + SC_TOTERR, c->scal + SC_TOTERR,
                               sizeof(double), hipMemcpyDeviceToHost, c->stream));
  QDEQ_HIP(c, hipMemcpyAsync(c->flags_host, c->flags, sizeof(int),
                             hipMemcpyDeviceToHost, c->stream));
  QDEQ_HIP(c, hipEventRecord(r.ev, c->stream));
  return QDEQ_OK;
}

// *ready = 1 when the last readback has landed.
static int rk_readback_ready(qdeq_ctx *c, bool block, int *ready) {
  RkMachine &r = c->rk;
  if (block) {
    QDEQ_HIP(c, hipEventSynchronize(r.ev));
    *ready = 1;
    return QDEQ_OK;
  }
  const hipError_t e = hipEventQuery(r.ev);
  if (e == hipErrorNotReady) {
    *ready = 0;
    return QDEQ_OK;
  }
  QDEQ_HIP(c, e);
  *ready = 1;
  return QDEQ_OK;
}

// Work may be launched ahead of a decision the host has not read yet for a
// single simulation launching kernel by kernel (FP64, start decided on the
// device, no candidate start, no fault-update timing) with nothing else
// launched ahead.  QDEQ_SPECULATE=0 turns it off.
static bool spec_base_allowed(const qdeq_ctx *c) {
  return !c->spec_off && launch_direct(c) && !stage_graphs(c) && device_start(c) &&
         c->cfg.precision == QDEQ_PREC_FP64 && c->timing != QDEQ_TIME_FAULT &&
         c->proj_slots == 0 && (c->guess == nullptr || c->guess_mode == QDEQ_GUESS_OFF) &&
         !c->spec_pending;
}
// The next stage behind the current solve: its V-cycles in flight, started
// on the device from the plain start.
static bool spec_allowed(const qdeq_ctx *c) {
  return spec_base_allowed(c) && c->mg.phase == 2 && c->mg.dev_start && !c->mg.cand;
}

// Whether the accepted attempt's commit and final stage-1 solve may be
// queued behind the combination, before the host has read the embedded
// error: the same single-simulation conditions, no solve in flight.
static bool accept_spec_allowed(const qdeq_ctx *c) {
  return spec_base_allowed(c) && c->mg.phase == 0;
}

// The step size an accepted step hands to the next one (odeSolver.cpp:1621-1632)
static double step_finish_dt(const qdeq_ctx *c) {
  const qdeq_config &p = c->cfg;
  if (c->totErr != 0.0)
    return qdeq_compute_step_size(c->it, c->dt, p.mindt, p.maxdt, c->totErr, p.totTol,
                                  c->errA);
  return c->dt;
}

// Whether stage 2 of the next step may be queued behind an accepted step's
// final stage-1 solve: the next step runs within this request and nothing is
// recorded between the steps (per-step outputs and events read the state
// there).
static bool spec_next_allowed(const qdeq_ctx *c) {
  return c->rk.steps_left > 1 && c->t < c->cfg.tmax && !c->out && !c->ev_on &&
         spec_allowed(c);
}

// Advances the machine as far as it can; *done = 1 when the requested steps
// (and the final flag check) are complete.  With block, every wait blocks.
static int rk_advance(qdeq_ctx *c, bool block, int *done) {
  RkMachine &r = c->rk;
  const qdeq_config &p = c->cfg;
  *done = 0;
  int st, ready;
  for (;;) {
    switch (r.phase) {
      case RK_FIRST:
        r.s = 1;
        r.final_stage = false;
        r.phase = RK_STAGE_BEGIN;
        break;
      case RK_STEP_BEGIN:
        // while (it < maxNumSteps && t_elapsed < tmax) (odeSolver.cpp:1462)
        if (r.steps_left <= 0 || !(c->t < p.tmax)) {
          if (r.spec)
            return set_error(c, QDEQ_ERR_STATE, "a stage is queued for a step that does not run");
          if ((st = rk_readback(c, false))) return st;
          r.phase = RK_WAIT_END;
          break;
        }
        c->it = c->it + 1;
        r.attempts = 0;
        r.phase = RK_ATTEMPT_BEGIN;
        break;
      case RK_ATTEMPT_BEGIN:  // odeSolver.cpp:1466-1604
        r.attempts = r.attempts + 1;
        if (c->t + c->dt > p.tmax) c->dt = p.tmax - c->t;
        if (r.spec) {
          // stage 2 already queued behind the last step's final solve, with
          // the dt computed there
          if (c->dt != r.spec_dt)
            return set_error(c, QDEQ_ERR_STATE, "step size differs from the queued stage's");
        } else {
          launch_zero_stage_rates(c, c->t, c->dt);
        }
        r.s = 2;
        r.final_stage = false;
        r.phase = RK_STAGE_BEGIN;
        break;
      case RK_STAGE_BEGIN:  // boundary data, then the elliptic solve
        c->solve_key = (r.final_stage ? 8 : 0) + r.s;
        if (r.spec) {
          // launched ahead behind the previous solve, which converged with
          // the work already queued: take it over
          c->mg = c->mg_next;
          c->mb_page = c->mg.page;
          c->spec_pending = false;
          c->spec_taken++;
          if (r.s == 2 && !r.final_stage) c->spec_taken_next++;
          r.spec = false;
          r.staged = false;
          r.phase = RK_SOLVING;
          break;
        }
        c->mb_page = 0;
        if (stage_graphs(c)) {
          // the stage's state (or the previous stage's fault update with
          // it), start and first V-cycles as one graph replay
          const int kind = r.staged ? 1 : 0;
          r.staged = false;
          if ((st = mg_begin_stage(c, kind, r.s))) return st;
          r.phase = RK_SOLVING;
          break;
        }
        if (!r.staged) launch_stage(c, r.s);
        r.staged = false;
        if ((st = guess_before_solve(c))) return st;
        if ((st = proj_before_solve(c))) return st;
        if ((st = mg_begin(c))) return st;
        r.phase = RK_SOLVING;
        break;
      case RK_SOLVING: {
        if (!r.spec && c->rk_ready && !r.final_stage && r.s < 6 && spec_allowed(c)) {
          // stage s + 1 queued behind this solve: the fault update of stage
          // s with stage s + 1's state, its start and first batch, all
          // gated on this solve's outcome
          launch_update_stage(c, r.s, c->gate);
          const int key = c->solve_key;
          c->solve_key = r.s + 1;
          st = mg_begin_spec(c);
          c->solve_key = key;
          if (st) return st;
          r.spec = true;
        } else if (!r.spec && c->rk_ready && r.final_stage && spec_next_allowed(c)) {
          // the next step's first attempt queued behind this solve: this
          // stage's fault update, the attempt's zeroed rates and stage 2's
          // state, start and first batch, with the dt the step will hand on
          double dt = step_finish_dt(c);
          if (c->t + dt > p.tmax) dt = p.tmax - c->t;
          launch_update_next(c, c->t, dt, c->gate);
          const int key = c->solve_key;
          c->solve_key = 2;
          st = mg_begin_spec(c);
          c->solve_key = key;
          if (st) return st;
          r.spec = true;
          r.spec_dt = dt;
        }
        int sdone = 0;
        st = mg_poll(c, block, &sdone);
        // a continuation of this solve abandoned the stage launched ahead
        if (r.spec && !c->spec_pending) {
          r.spec = false;
          c->spec_abandoned++;
        }
        if (st) {
          r.spec = false;
          c->spec_pending = false;
          return st;
        }
        if (!sdone) return QDEQ_OK;
        if ((st = guess_after_solve(c))) return st;
        if ((st = proj_after_solve(c))) return st;
        if (c->timing == QDEQ_TIME_FAULT && (c->timed_seq++ % 8) == 0) {
          QDEQ_HIP(c, hipEventRecord(c->ev_t[0][0], c->stream));
          launch_fault_update(c, r.s);
          QDEQ_HIP(c, hipEventRecord(c->ev_t[0][1], c->stream));
          QDEQ_HIP(c, hipEventSynchronize(c->ev_t[0][1]));
          float ms = 0.0f;
          if (hipEventElapsedTime(&ms, c->ev_t[0][0], c->ev_t[0][1]) == hipSuccess) {
            c->timed_ms += ms;
            c->timed_calls += 1;
          }
        } else if (c->rk_ready && !r.final_stage && r.s < 6) {
          // the fault update and stage s + 1's state: one launch, or the
          // first node of stage s + 1's graph, or already queued with
          // stage s + 1 launched ahead
          if (!r.spec) {
            if (!stage_graphs(c)) launch_update_stage(c, r.s);
            r.staged = true;
          }
        } else if (!r.spec) {
          launch_fault_update(c, r.s);
        }
        if (c->rk_ready && !r.final_stage) {
          if (r.s < 6) {
            r.s++;
            r.phase = RK_STAGE_BEGIN;
          } else {
            launch_combine(c);
            if ((st = rk_readback(c, true))) return st;
            if (accept_spec_allowed(c)) {
              // the accepted case queued behind the readback, gated on the
              // combination's own accept decision (gate[7]): the commit, the final
              // stage 1's state, and that solve's start and first batch
              launch_commit(c, c->t + c->dt, c->gate);
              launch_stage(c, 1, c->gate);
              const int key = c->solve_key;
              c->solve_key = 8 + 1;
              st = mg_begin_spec(c, 7);
              c->solve_key = key;
              if (st) return st;
              r.spec = true;
            }
            r.phase = RK_WAIT_COMBINE;
          }
        } else if (!c->rk_ready) {
          if ((st = rk_readback(c, false))) return st;
          r.phase = RK_WAIT_FIRST;
        } else {
          r.phase = RK_STEP_FINISH;
        }
        break;
      }
      case RK_WAIT_FIRST:
        if ((st = rk_readback_ready(c, block, &ready))) return st;
        if (!ready) return QDEQ_OK;
        if ((st = check_flags(c))) return st;
        c->rk_ready = true;
        r.phase = RK_STEP_BEGIN;
        break;
      case RK_WAIT_COMBINE: {
        if ((st = rk_readback_ready(c, block, &ready))) return st;
        if (!ready) return QDEQ_OK;
        if ((st = check_flags(c))) return st;
        c->totErr = c->scal_host[SC_TOTERR];
        bool accept = c->totErr < p.totTol;
        if (r.spec && !accept) {
          // rejected on the device too: the work queued for the accepted
          // case did nothing (its start abandoned itself)
          r.spec = false;
          c->spec_pending = false;
          c->spec_abandoned++;
        }
        if (!accept) {
          c->dt = qdeq_compute_step_size(c->it, c->dt, p.mindt, p.maxdt, c->totErr,
                                         p.totTol, c->errA);
          accept = p.mindt == c->dt || r.attempts >= 100;
        }
        if (!accept) {
          r.phase = RK_ATTEMPT_BEGIN;
          break;
        }
        c->attempts = r.attempts;
        c->t = c->t + c->dt;  // odeSolver.cpp:1605-1619
        // already queued (with the final stage 1) when the device accepted
        if (!r.spec) launch_commit(c, c->t);
        r.s = 1;
        r.final_stage = true;
        r.phase = RK_STAGE_BEGIN;
        break;
      }
      case RK_STEP_FINISH: {
        const double newDeltaT = step_finish_dt(c);
        c->errA[1] = c->errA[0];
        c->errA[0] = c->totErr;
        c->dt = newDeltaT;
        // per-step record (odeSolver.cpp:1634-1652): events, output file
        if (c->ev_on) events_after_step(c);
        if ((st = output_after_step(c))) return st;
        r.steps_left--;
        r.phase = RK_STEP_BEGIN;
        break;
      }
      case RK_WAIT_END:
        if ((st = rk_readback_ready(c, block, &ready))) return st;
        if (!ready) return QDEQ_OK;
        if ((st = check_flags(c))) return st;
        r.phase = RK_DONE;
        *done = 1;
        return QDEQ_OK;
      case RK_DONE:
        *done = 1;
        return QDEQ_OK;
      default:
        return set_error(c, QDEQ_ERR_STATE, "RK43 driver has no steps requested");
    }
  }
}

// Requests nsteps more accepted steps (and the t = 0 stage first if needed).
static int rk_request(qdeq_ctx *c, int nsteps) {
  if (c->mg.phase != 0 ||
      (c->rk.phase != RK_IDLE && c->rk.phase != RK_DONE))
    return set_error(c, QDEQ_ERR_STATE, "RK43 driver is busy");
  c->rk.steps_left = nsteps;
  c->rk.phase = c->rk_ready ? RK_STEP_BEGIN : RK_FIRST;
  return QDEQ_OK;
}

int qdeq_rk43_step(qdeq_ctx *c, int nsteps, qdeq_step_info *info) {
  if (!c || nsteps < 0) return set_error(c, QDEQ_ERR_ARG, "bad argument");
  int st = rk_request(c, nsteps);
  if (st != QDEQ_OK) return st;
  int done = 0;
  while (!done) {
    if ((st = rk_advance(c, true, &done))) {
      c->rk.phase = RK_IDLE;
      c->mg.phase = 0;
      return st;
    }
  }
  c->rk.phase = RK_IDLE;
  if (info) return qdeq_rk43_get_state(c, info, nullptr, nullptr, nullptr,
                                       nullptr, nullptr, nullptr);
  return QDEQ_OK;
}

int qdeq_folded_stop_tests(qdeq_ctx *c, long long *folded) {
  if (!c || !folded) return set_error(c, QDEQ_ERR_ARG, "null argument");
  *folded = c->folded_tests;
  return QDEQ_OK;
}

int qdeq_launch_ahead_stats(qdeq_ctx *c, long long *taken, long long *abandoned,
                            long long *taken_next_step) {
  if (!c) return set_error(c, QDEQ_ERR_ARG, "null argument");
  if (taken) *taken = c->spec_taken;
  if (abandoned) *abandoned = c->spec_abandoned;
  if (taken_next_step) *taken_next_step = c->spec_taken_next;
  return QDEQ_OK;
}

int qdeq_ensemble_step(qdeq_ctx **ctxs, int n, int nsteps, int *status) {
  if (!ctxs || n < 1 || nsteps < 0)
    return set_error(nullptr, QDEQ_ERR_ARG, "bad argument");
  std::vector<int> done((size_t)n, 0);
  int st_all = QDEQ_OK;
  for (int i = 0; i < n; i++)
    if (ctxs[i]) ctxs[i]->in_ensemble = true;
  for (int i = 0; i < n; i++) {
    int st = ctxs[i] ? rk_request(ctxs[i], nsteps) : QDEQ_ERR_ARG;
    if (status) status[i] = st;
    if (st != QDEQ_OK) {
      done[(size_t)i] = 1;
      st_all = st;
    }
  }
  // round robin: every member advances as far as it can without waiting
  int left = 0;
  for (int i = 0; i < n; i++) left += done[(size_t)i] ? 0 : 1;
  while (left > 0) {
    for (int i = 0; i < n; i++) {
      if (done[(size_t)i]) continue;
      qdeq_ctx *c = ctxs[i];
      int d = 0;
      const int st = rk_advance(c, false, &d);
      if (st != QDEQ_OK || d) {
        if (st != QDEQ_OK) {
          // let the member's queued work drain before it is reused
          (void)hipStreamSynchronize(c->stream);
          c->mg.phase = 0;
          st_all = st;
        }
        c->rk.phase = RK_IDLE;
        if (status) status[i] = st;
        done[(size_t)i] = 1;
        left--;
      }
    }
  }
  for (int i = 0; i < n; i++)
    if (ctxs[i]) ctxs[i]->in_ensemble = false;
  return st_all;
}

int qdeq_rk43_get_state(qdeq_ctx *c, qdeq_step_info *info, double *U,
                        double *slip, double *psi, double *V, double *dpsi,
                        double *Ts) {
  if (!c) return set_error(c, QDEQ_ERR_ARG, "null argument");
  const size_t nz = (size_t)c->nz, ng = (size_t)c->nx * c->nz;
  FaultBufs &f = c->fb;
  if (info) {
    launch_maxima(c);
    QDEQ_HIP(c, hipMemcpyAsync(c->scal_host + SC_VMAX, c->scal + SC_VMAX,
                               2 * sizeof(double), hipMemcpyDeviceToHost,
                               c->stream));
  }
  if (U) {
    QDEQ_HIP(c, hipMemsetAsync(c->stage_U, 0, ng * sizeof(double), c->stream));
    launch_unpack(c, c->stage_U);
    launch_extend_crew(c, c->stage_U);
    QDEQ_HIP(c, hipMemcpyAsync(U, c->stage_U, ng * sizeof(double),
                               hipMemcpyDeviceToHost, c->stream));
  }
  struct {
    double *dst;
    const double *src;
  } dn[] = {{slip, f.slip}, {psi, f.psi}, {V, f.FV[1]}, {dpsi, f.Fd[1]}, {Ts, f.Ts}};
  for (auto &d : dn)
    if (d.dst)
      QDEQ_HIP(c, hipMemcpyAsync(d.dst, d.src, nz * sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream));
  QDEQ_HIP(c, hipGetLastError());
  QDEQ_HIP(c, hipStreamSynchronize(c->stream));
  if (info) {
    info->t = c->t;
    info->dt = c->dt;
    info->totErr = c->totErr;
    info->errA0 = c->errA[0];
    info->errA1 = c->errA[1];
    info->V_max = c->scal_host[SC_VMAX];
    info->Ts_max = c->scal_host[SC_TSMAX];
    info->it = c->it;
    info->attempts = c->attempts;
    info->vcycles = c->vcycles;
    info->solves = c->solves;
    info->ir_sweeps = c->ir_sweeps;
  }
  return QDEQ_OK;
}

}  // extern "C"
