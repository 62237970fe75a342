// guess.hip -- the learned initial guess of the RK43 stage solves: network
// object, pair harvesting, training and the per-solve prediction.  No
// reference counterpart; the reference's solve starts from the previous
// solution (odeSolver.cpp:652-720), which remains the fallback here.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "nn.h"
#include "qdeq_guess.h"
#include "qdeq_internal.h"

using qdeq::nn::bf16;

struct qdeq_guess {
  qdeq_guess_config cfg;
  hipStream_t stream = nullptr;
  std::string err;
  int nk, f, nc, nout, noutp, C, T, P0, F0, H;
  // parameter tensors in one FP32 allocation (W4 padded to noutp rows)
  long long off[5], cnt[5], ntot;
  float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;
  bf16 *W[4] = {}, *WT[4] = {};  // BF16 copies [out x in] and [in x out]
  long long adam_t = 0;
  // harvested pairs
  float *xs = nullptr, *ts = nullptr;
  double *scales = nullptr, *mags = nullptr;  // max|ds| and max|boundary data|
  int count = 0;
  // training scratch for Mb rows
  int tb_M = 0;
  float *X = nullptr, *Tg = nullptr, *Z = nullptr, *Y = nullptr, *dZ0 = nullptr;
  bf16 *Hb[4] = {}, *HbT[4] = {};     // H0..H3 and transposes
  bf16 *dY = nullptr, *dYT = nullptr;
  bf16 *dZb[2] = {}, *dZbT[2] = {};
  float *split = nullptr;
  size_t split_n = 0;
  double *mse_part = nullptr, *loss = nullptr;
  int *idx = nullptr;
  int idx_n = 0;
};

namespace qdeq {

namespace {

int gset_error(qdeq_guess *g, int code, const std::string &msg) {
  if (g) g->err = msg;
  return code;
}

#define GHIP(g, call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return gset_error((g), QDEQ_ERR_HIP,                                   \
                        std::string(#call) + ": " + hipGetErrorString(e_));  \
  } while (0)

template <typename T>
hipError_t dalloc(T **p, size_t n) {
  hipError_t e = hipMalloc(p, n * sizeof(T));
  if (e != hipSuccess) return e;
  return hipMemset(*p, 0, n * sizeof(T));
}

long long round_up(long long a, long long b) { return (a + b - 1) / b * b; }

// Tensor t of the parameter vector as a matrix: rows x cols ([out x in]).
void tensor_shape(const qdeq_guess *g, int t, int *rows, int *cols, int *rows_pad) {
  const int r[5] = {g->C, g->H, g->H, g->H, g->nout};
  const int c[5] = {g->T, g->F0, g->H, g->H, g->H};
  *rows = r[t];
  *cols = c[t];
  *rows_pad = t == 4 ? g->noutp : r[t];
}

// BF16 weight copies from the FP32 master weights.
void refresh_bf16(qdeq_guess *g, hipStream_t st) {
  for (int l = 0; l < 4; l++) {
    int rows, cols, rp;
    tensor_shape(g, l + 1, &rows, &cols, &rp);
    nn::epilogue(st, g->w + g->off[l + 1], rp, cols, nn::EPI_PLAIN, nullptr,
                 g->W[l], g->WT[l], nullptr);
  }
}

int ensure_train_bufs(qdeq_guess *g, int Mb) {
  if (g->tb_M >= Mb) return QDEQ_OK;
  void *old[] = {g->X, g->Tg, g->Z, g->Y, g->dZ0, g->dY, g->dYT, g->split,
                 g->mse_part, g->loss};
  for (void *p : old) (void)hipFree(p);
  for (int i = 0; i < 4; i++) {
    (void)hipFree(g->Hb[i]);
    (void)hipFree(g->HbT[i]);
  }
  for (int i = 0; i < 2; i++) {
    (void)hipFree(g->dZb[i]);
    (void)hipFree(g->dZbT[i]);
  }
  g->tb_M = 0;
  const size_t M = (size_t)Mb, wmax = (size_t)std::max(g->F0, g->H);
  GHIP(g, dalloc(&g->X, M * g->nk));
  GHIP(g, dalloc(&g->Tg, M * g->noutp));
  GHIP(g, dalloc(&g->Z, M * wmax));
  GHIP(g, dalloc(&g->Y, M * g->noutp));
  GHIP(g, dalloc(&g->dZ0, M * g->F0));
  GHIP(g, dalloc(&g->dY, M * g->noutp));
  GHIP(g, dalloc(&g->dYT, M * g->noutp));
  for (int i = 0; i < 4; i++) {
    const size_t w = i == 0 ? g->F0 : g->H;
    GHIP(g, dalloc(&g->Hb[i], M * w));
    GHIP(g, dalloc(&g->HbT[i], M * w));
  }
  for (int i = 0; i < 2; i++) {
    GHIP(g, dalloc(&g->dZb[i], M * g->H));
    GHIP(g, dalloc(&g->dZbT[i], M * g->H));
  }
  g->split_n = (size_t)64 * M * wmax;
  GHIP(g, dalloc(&g->split, g->split_n));
  GHIP(g, dalloc(&g->mse_part, (size_t)nn::MSE_BLOCKS));
  GHIP(g, dalloc(&g->loss, (size_t)1 << 16));
  g->tb_M = Mb;
  return QDEQ_OK;
}

int gemm(qdeq_guess *g, hipStream_t st, const bf16 *A, int lda, const bf16 *B,
         int ldb, float *C, int M, int N, int K) {
  if (nn::gemm_nt(st, nn::GEMM_CFG_WIDE, A, lda, B, ldb, C, N, M, N, K, g->split,
                  g->split_n, 512))
    return gset_error(g, QDEQ_ERR_ARG, "GEMM shape not covered by the tiles");
  return QDEQ_OK;
}

// Forward pass on rows X[0..Mb) (already in g->X); Y = output (FP32).
int forward(qdeq_guess *g, hipStream_t st, int Mb) {
  nn::conv_forward(st, g->X, Mb, g->nk, g->w + g->off[0], g->C, g->T, g->f,
                   g->P0, g->Hb[0], g->HbT[0]);
  for (int l = 0; l < 3; l++) {
    const int in = l == 0 ? g->F0 : g->H;
    int st_ = gemm(g, st, g->Hb[l], in, g->W[l], in, g->Z, Mb, g->H, in);
    if (st_) return st_;
    nn::epilogue(st, g->Z, Mb, g->H, nn::EPI_RELU, nullptr, g->Hb[l + 1],
                 g->HbT[l + 1], nullptr);
  }
  return gemm(g, st, g->Hb[3], g->H, g->W[3], g->H, g->Y, Mb, g->noutp, g->H);
}

// Backward pass after forward + mse (g->Y holds dloss/dY): gradients in g->g.
int backward(qdeq_guess *g, hipStream_t st, int Mb) {
  int s;
  nn::epilogue(st, g->Y, Mb, g->noutp, nn::EPI_PLAIN, nullptr, g->dY, g->dYT,
               nullptr);
  // W4: [noutp x H] = dY^T [noutp x Mb] * H3^T [H x Mb]^T
  if ((s = gemm(g, st, g->dYT, Mb, g->HbT[3], Mb, g->g + g->off[4], g->noutp,
                g->H, Mb)))
    return s;
  // dH3 [Mb x H] = dY [Mb x noutp] * W4^T [H x noutp]^T
  if ((s = gemm(g, st, g->dY, g->noutp, g->WT[3], g->noutp, g->Z, Mb, g->H,
                g->noutp)))
    return s;
  const bf16 *dz = nullptr, *dzT = nullptr;
  for (int l = 2; l >= 0; l--) {
    bf16 *o = g->dZb[l & 1], *oT = g->dZbT[l & 1];
    // dZ_{l+1} = dH_{l+1} * (H_{l+1} > 0)
    nn::epilogue(st, g->Z, Mb, g->H, nn::EPI_MASK, g->Hb[l + 1], o, oT, nullptr);
    dz = o;
    dzT = oT;
    const int in = l == 0 ? g->F0 : g->H;
    // W_{l+1}: [H x in] = dZ^T [H x Mb] * H_l^T [in x Mb]^T
    if ((s = gemm(g, st, dzT, Mb, g->HbT[l], Mb, g->g + g->off[l + 1], g->H, in, Mb)))
      return s;
    // dH_l [Mb x in] = dZ [Mb x H] * W_{l+1}^T [in x H]^T
    if ((s = gemm(g, st, dz, g->H, g->WT[l], g->H, g->Z, Mb, in, g->H))) return s;
  }
  // encoder: dZ0 = dH0 * (H0 > 0) (FP32), then the conv weight gradient
  nn::epilogue(st, g->Z, Mb, g->F0, nn::EPI_MASK, g->Hb[0], nullptr, nullptr,
               g->dZ0);
  nn::conv_backward(st, g->X, Mb, g->nk, g->dZ0, g->C, g->T, g->f, g->P0,
                    g->g + g->off[0]);
  return QDEQ_OK;
}

// Copies host rows into the padded device matrix (rows >= B zero).
int upload_rows(qdeq_guess *g, float *dst, int Mb, int ld, const float *src,
                int B, int cols) {
  GHIP(g, hipMemsetAsync(dst, 0, (size_t)Mb * ld * sizeof(float), g->stream));
  GHIP(g, hipMemcpy2DAsync(dst, (size_t)ld * sizeof(float), src,
                           (size_t)cols * sizeof(float), (size_t)cols * sizeof(float),
                           (size_t)B, hipMemcpyHostToDevice, g->stream));
  return QDEQ_OK;
}

// ---- per-solve kernels of a context ----

constexpr int PREP_THREADS = 1024;
constexpr double kMinRelIncrement = 1e-7;

// ds = fault column now - previous; s = max|ds|; dR = remote change (row 0);
// mag = max|boundary data| (fault column, remote value of row 0).
// x = ds / s (FP32).  With `before`, the interior coarse samples of u are
// saved for the harvested target.
__global__ __launch_bounds__(PREP_THREADS) void k_guess_prep(
    const double *__restrict__ u, int N, int K, int P,
    const double *__restrict__ gprev, float *__restrict__ x,
    double *__restrict__ gs, double *__restrict__ before, int f, int nc) {
  __shared__ double red[PREP_THREADS], redm[PREP_THREADS];
  double mx = 0.0, mm = 0.0;
  for (int j = threadIdx.x; j < N; j += PREP_THREADS) {
    const double ub = u[(size_t)j * P];
    mx = fmax(mx, fabs(ub - gprev[j]));
    mm = fmax(mm, fabs(ub));
  }
  red[threadIdx.x] = mx;
  redm[threadIdx.x] = mm;
  __syncthreads();
  for (int s = PREP_THREADS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
      redm[threadIdx.x] = fmax(redm[threadIdx.x], redm[threadIdx.x + s]);
    }
    __syncthreads();
  }
  const double smax = red[0];
  for (int j = threadIdx.x; j < N; j += PREP_THREADS) {
    const double ds = u[(size_t)j * P] - gprev[j];
    x[j] = smax > 0.0 ? (float)(ds / smax) : 0.0f;
  }
  if (threadIdx.x == 0) {
    gs[0] = smax;
    gs[1] = u[K - 1] - gprev[N];
    gs[3] = fmax(redm[0], fabs(u[K - 1]));
  }
  if (before)
    for (int i = threadIdx.x; i < nc * nc; i += PREP_THREADS) {
      const int J = i / nc, Kc = i % nc;
      before[i] = u[(size_t)(f * J) * P + f * Kc];
    }
}

// Target of pair `slot`: (u_after - before - dR W_R) / s at the interior
// coarse points, x[f J] on the fault column, 0 on the remote column.
__global__ void k_guess_harvest(const double *__restrict__ u, int P,
                                const double *__restrict__ before,
                                const double *__restrict__ wr,
                                const double *__restrict__ gs,
                                const float *__restrict__ x, int f, int nc,
                                float *__restrict__ t, double *__restrict__ scale,
                                double *__restrict__ mag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc * nc) return;
  const int J = i / nc, Kc = i % nc;
  const double s = gs[0], dR = gs[1];
  float v;
  if (Kc == 0)
    v = x[f * J];
  else if (Kc == nc - 1)
    v = 0.0f;
  else {
    const size_t c = (size_t)(f * J) * P + f * Kc;
    v = s > 0.0 ? (float)((u[c] - before[i] - dR * wr[c]) / s) : 0.0f;
  }
  t[i] = v;
  if (i == 0) {
    *scale = s;
    *mag = gs[3];
  }
}

// Boundary data of this solve become the reference of the next one.
__global__ void k_guess_commit(const double *__restrict__ u, int N, int K, int P,
                               double *__restrict__ gprev) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < N;
       j += gridDim.x * blockDim.x)
    gprev[j] = u[(size_t)j * P];
  if (blockIdx.x == 0 && threadIdx.x == 0) gprev[N] = u[K - 1];
}

// cand = u + s * factor * P(y) + dR * W_R on interior columns, u elsewhere.
// P: bilinear from the coarse grid (every f-th node); on the first coarse
// column the exact fine-row boundary change x[j] is used.
__global__ void k_guess_apply(const double *__restrict__ u,
                              const double *__restrict__ wr,
                              const float *__restrict__ y,
                              const float *__restrict__ x,
                              const double *__restrict__ gs, int N, int K, int P,
                              int f, int nc, double *__restrict__ cand) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (k >= P || j >= N) return;
  const size_t c = (size_t)j * P + k;
  if (k == 0 || k >= K - 1) {
    cand[c] = u[c];
    return;
  }
  const double s = gs[0] * gs[2], dR = gs[1];
  const int J0 = min(j / f, nc - 2), K0 = min(k / f, nc - 2);
  const double wj = (double)(j - f * J0) / f, wk = (double)(k - f * K0) / f;
  const float *r0 = y + (size_t)J0 * nc + K0;
  const float *r1 = r0 + nc;
  const double left = K0 == 0 ? (double)x[j]
                              : (1.0 - wj) * (double)r0[0] + wj * (double)r1[0];
  const double right = (1.0 - wj) * (double)r0[1] + wj * (double)r1[1];
  const double val = (1.0 - wk) * left + wk * right;
  cand[c] = u[c] + s * val + dR * wr[c];
}

__global__ void k_unit_remote(double *__restrict__ u, int N, int K, int P) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (k >= P || j >= N) return;
  u[(size_t)j * P + k] = k == K - 1 ? 1.0 : 0.0;
}

}  // namespace

// Enqueues the inference chain + candidate + its residual norm on c->stream.
static void launch_predict_chain(qdeq_ctx *c) {
  qdeq_guess *g = c->guess;
  const LevelDims &d = c->lv[0];
  hipStream_t st = c->stream;
  hipLaunchKernelGGL(k_guess_prep, dim3(1), dim3(PREP_THREADS), 0, st, c->d[0].u,
                     d.N, d.K, d.P, c->gprev, c->gx, c->gs, (double *)nullptr,
                     g->f, g->nc);
  nn::conv_forward(st, c->gx, 16, g->nk, g->w + g->off[0], g->C, g->T, g->f,
                   g->P0, c->gh[0], nullptr);
  for (int l = 0; l < 3; l++) {
    const int in = l == 0 ? g->F0 : g->H;
    nn::gemm_nt(st, nn::GEMM_CFG_SKINNY, c->gh[l], in, g->W[l], in, c->gz, g->H,
                16, g->H, in, c->gsplit, c->gsplit_n, 64);
    nn::epilogue(st, c->gz, 16, g->H, nn::EPI_RELU, nullptr, c->gh[l + 1], nullptr,
                 nullptr);
  }
  nn::gemm_nt(st, nn::GEMM_CFG_SKINNY, c->gh[3], g->H, g->W[3], g->H, c->gy,
              g->noutp, 16, g->noutp, g->H, c->gsplit, c->gsplit_n, 64);
  const dim3 blk(256);
  const dim3 grd((d.P + 255) / 256, d.N);
  hipLaunchKernelGGL(k_guess_apply, grd, blk, 0, st, c->d[0].u, c->wr, c->gy, c->gx,
                     c->gs, d.N, d.K, d.P, g->f, g->nc, c->d[0].r);
  launch_resnorm_of(c, c->d[0].r, 2 * kNormCap, SC_NORM_PRED);
}

int guess_before_solve(qdeq_ctx *c) {
  c->cand_pending = false;
  if (!c->guess || c->guess_mode == QDEQ_GUESS_OFF || !c->gprev_valid)
    return QDEQ_OK;
  qdeq_guess *g = c->guess;
  const LevelDims &d = c->lv[0];
  if (c->guess_mode == QDEQ_GUESS_HARVEST) {
    if (g->count >= g->cfg.capacity) return QDEQ_OK;
    if (c->g_seen++ % g->cfg.stride != 0) return QDEQ_OK;
    hipLaunchKernelGGL(k_guess_prep, dim3(1), dim3(PREP_THREADS), 0, c->stream,
                       c->d[0].u, d.N, d.K, d.P, c->gprev,
                       g->xs + (size_t)g->count * g->nk, c->gs, c->gbefore, g->f,
                       g->nc);
    c->harvest_pending = true;
    return QDEQ_OK;
  }
  if (launch_direct(c)) {
    mbox_arm(c, SC_NORM_PRED);
    launch_predict_chain(c);
    QDEQ_HIP(c, hipGetLastError());
    c->cand_pending = true;
    c->g_predicted++;
    return QDEQ_OK;
  }
  if (!c->g_pred) {
    hipGraph_t gr = nullptr;
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      launch_predict_chain(c);
      e = hipStreamEndCapture(c->stream, &gr);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&c->g_pred, gr, nullptr, nullptr, 0);
    if (gr) (void)hipGraphDestroy(gr);
    if (e != hipSuccess)
      return set_error(c, QDEQ_ERR_HIP,
                       std::string("guess graph capture: ") + hipGetErrorString(e));
  }
  mbox_arm(c, SC_NORM_PRED);
  QDEQ_HIP(c, hipGraphLaunch(c->g_pred, c->stream));
  c->cand_pending = true;
  c->g_predicted++;
  return QDEQ_OK;
}

int guess_after_solve(qdeq_ctx *c) {
  if (!c->guess || c->guess_mode == QDEQ_GUESS_OFF) return QDEQ_OK;
  qdeq_guess *g = c->guess;
  const LevelDims &d = c->lv[0];
  if (c->harvest_pending) {
    const int slot = g->count;
    hipLaunchKernelGGL(k_guess_harvest, dim3((g->nout + 255) / 256), dim3(256), 0,
                       c->stream, c->d[0].u, d.P, c->gbefore, c->wr, c->gs,
                       g->xs + (size_t)slot * g->nk, g->f, g->nc,
                       g->ts + (size_t)slot * g->noutp, g->scales + slot,
                       g->mags + slot);
    g->count++;
    c->g_harvested++;
    c->harvest_pending = false;
  }
  hipLaunchKernelGGL(k_guess_commit, dim3((d.N + 255) / 256), dim3(256), 0,
                     c->stream, c->d[0].u, d.N, d.K, d.P, c->gprev);
  QDEQ_HIP(c, hipGetLastError());
  c->gprev_valid = true;
  return QDEQ_OK;
}

void guess_reset(qdeq_ctx *c) {
  c->gprev_valid = false;
  c->cand_pending = false;
  c->harvest_pending = false;
}

void guess_free_ctx(qdeq_ctx *c) {
  if (c->g_pred) (void)hipGraphExecDestroy(c->g_pred);
  c->g_pred = nullptr;
  void *p[] = {c->wr, c->gx, c->gz, c->gy, c->gsplit, c->gprev, c->gbefore, c->gs};
  for (void *q : p) (void)hipFree(q);
  for (auto &h : c->gh) (void)hipFree(h);
  c->wr = nullptr;
  c->gx = nullptr;
  c->gz = nullptr;
  c->gy = nullptr;
  c->gsplit = nullptr;
  c->gprev = nullptr;
  c->gbefore = nullptr;
  c->gs = nullptr;
  for (auto &h : c->gh) h = nullptr;
  c->guess = nullptr;
}

// W_R: solution for unit remote data and zero fault data, by one solve to
// a relative residual of 1e-12.
int solve_unit_remote(qdeq_ctx *c, double *dst) {
  const LevelDims &d = c->lv[0];
  const size_t n = (size_t)d.N * d.P;
  if ((size_t)2 * c->nx * c->nz < n)
    return set_error(c, QDEQ_ERR_STATE, "staging buffer too small for W_R");
  QDEQ_HIP(c, hipMemcpyAsync(c->stage_U, c->d[0].u, n * sizeof(double),
                             hipMemcpyDeviceToDevice, c->stream));
  hipLaunchKernelGGL(k_unit_remote, dim3((d.P + 255) / 256, d.N), dim3(256), 0,
                     c->stream, c->d[0].u, d.N, d.K, d.P);
  const qdeq_config saved = c->cfg;
  const long long sv = c->vcycles, ss = c->solves, si = c->ir_sweeps;
  c->cfg.tol_mode = QDEQ_TOL_REL;
  c->cfg.tol = 1e-12;
  c->cfg.max_vcycles = std::max(c->cfg.max_vcycles, 400);
  mg_destroy_graphs(c);
  int st = mg_solve(c, nullptr, nullptr, nullptr);
  c->cfg = saved;
  mg_destroy_graphs(c);
  c->vcycles = sv;
  c->solves = ss;
  c->ir_sweeps = si;
  if (st != QDEQ_OK) return st;
  QDEQ_HIP(c, hipMemcpyAsync(dst, c->d[0].u, n * sizeof(double),
                             hipMemcpyDeviceToDevice, c->stream));
  QDEQ_HIP(c, hipMemcpyAsync(c->d[0].u, c->stage_U, n * sizeof(double),
                             hipMemcpyDeviceToDevice, c->stream));
  QDEQ_HIP(c, hipStreamSynchronize(c->stream));
  return QDEQ_OK;
}

}  // namespace qdeq

using namespace qdeq;

extern "C" {

void qdeq_guess_default_config(qdeq_guess_config *c, int nk) {
  std::memset(c, 0, sizeof(*c));
  c->nk = nk;
  c->factor = 4;
  c->channels = 8;
  c->width = 8;
  c->hidden = 512;
  c->capacity = 20000;
  c->stride = 1;
  c->seed = 1234;
  c->device = 0;
}

void qdeq_train_default_config(qdeq_train_config *c) {
  std::memset(c, 0, sizeof(*c));
  c->epochs = 30;
  c->batch = 256;
  c->lr = 1e-3;
  c->lr_final = 3e-5;
  c->beta1 = 0.9;
  c->beta2 = 0.999;
  c->eps = 1e-8;
  c->seed = 7;
}

int qdeq_guess_create(const qdeq_guess_config *cfg, qdeq_guess **out) {
  if (!cfg || !out) return QDEQ_ERR_ARG;
  *out = nullptr;
  const int nk = cfg->nk, f = cfg->factor;
  if (nk < 33 || f < 2 || (nk - 1) % f || cfg->channels < 1 || cfg->width < f ||
      ((cfg->width - f) & 1) || cfg->hidden < 128 || cfg->hidden % 128 ||
      cfg->capacity < 0 || cfg->stride < 1 ||
      (((nk - 1) / f) * cfg->channels) % 128)
    return set_error(nullptr, QDEQ_ERR_ARG,
                     "guess config: need (nk-1) % factor == 0, width >= factor "
                     "with width-factor even, hidden % 128 == 0 and "
                     "(nk-1)/factor*channels % 128 == 0");
  if (hipSetDevice(cfg->device) != hipSuccess)
    return set_error(nullptr, QDEQ_ERR_HIP, "hipSetDevice failed");
  qdeq_guess *g = new qdeq_guess();
  g->cfg = *cfg;
  g->nk = nk;
  g->f = f;
  g->nc = (nk - 1) / f + 1;
  g->nout = g->nc * g->nc;
  g->noutp = (int)round_up(g->nout, 256);
  g->C = cfg->channels;
  g->T = cfg->width;
  g->P0 = (nk - 1) / f;
  g->F0 = g->P0 * g->C;
  g->H = cfg->hidden;
  long long o = 0;
  for (int t = 0; t < 5; t++) {
    int rows, cols, rp;
    tensor_shape(g, t, &rows, &cols, &rp);
    g->off[t] = o;
    g->cnt[t] = (long long)rp * cols;
    o += round_up(g->cnt[t], 64);
  }
  g->ntot = o;
  auto fail = [&](const char *what) {
    set_error(nullptr, QDEQ_ERR_NOMEM, std::string("guess allocation: ") + what);
    qdeq_guess_destroy(g);
    return QDEQ_ERR_NOMEM;
  };
  if (hipStreamCreate(&g->stream) != hipSuccess) return fail("stream");
  if (dalloc(&g->w, (size_t)o) || dalloc(&g->m, (size_t)o) || dalloc(&g->v, (size_t)o) ||
      dalloc(&g->g, (size_t)o))
    return fail("parameters");
  for (int l = 0; l < 4; l++)
    if (dalloc(&g->W[l], (size_t)g->cnt[l + 1]) || dalloc(&g->WT[l], (size_t)g->cnt[l + 1]))
      return fail("bf16 weights");
  const size_t cap = (size_t)std::max(cfg->capacity, 1);
  if (dalloc(&g->xs, cap * nk) || dalloc(&g->ts, cap * g->noutp) ||
      dalloc(&g->scales, cap) || dalloc(&g->mags, cap))
    return fail("pair store");
  // He-uniform initialisation, seeded (FP32 master weights)
  std::mt19937_64 rng(cfg->seed);
  std::vector<float> hw((size_t)o, 0.0f);
  for (int t = 0; t < 5; t++) {
    int rows, cols, rp;
    tensor_shape(g, t, &rows, &cols, &rp);
    const double bound = t == 4 ? std::sqrt(3.0 / cols) : std::sqrt(6.0 / cols);
    std::uniform_real_distribution<double> U(-bound, bound);
    for (long long i = 0; i < (long long)rows * cols; i++)
      hw[(size_t)(g->off[t] + i)] = (float)U(rng);
  }
  if (hipMemcpy(g->w, hw.data(), (size_t)o * sizeof(float), hipMemcpyHostToDevice) !=
      hipSuccess)
    return fail("upload");
  refresh_bf16(g, g->stream);
  if (hipStreamSynchronize(g->stream) != hipSuccess) return fail("init");
  *out = g;
  return QDEQ_OK;
}

int qdeq_guess_destroy(qdeq_guess *g) {
  if (!g) return QDEQ_OK;
  void *p[] = {g->w, g->m, g->v, g->g, g->xs, g->ts, g->scales, g->mags, g->X, g->Tg,
               g->Z, g->Y, g->dZ0, g->dY, g->dYT, g->split, g->mse_part,
               g->loss, g->idx};
  for (void *q : p) (void)hipFree(q);
  for (int i = 0; i < 4; i++) {
    (void)hipFree(g->W[i]);
    (void)hipFree(g->WT[i]);
    (void)hipFree(g->Hb[i]);
    (void)hipFree(g->HbT[i]);
  }
  for (int i = 0; i < 2; i++) {
    (void)hipFree(g->dZb[i]);
    (void)hipFree(g->dZbT[i]);
  }
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
  return QDEQ_OK;
}

const char *qdeq_guess_last_error(const qdeq_guess *g) {
  return g ? g->err.c_str() : "";
}

long long qdeq_guess_num_params(const qdeq_guess *g) {
  if (!g) return 0;
  long long n = 0;
  for (int t = 0; t < 5; t++) {
    int rows, cols, rp;
    tensor_shape(g, t, &rows, &cols, &rp);
    n += (long long)rows * cols;
  }
  return n;
}

int qdeq_guess_dims(const qdeq_guess *g, int *nout, int *f0) {
  if (!g) return QDEQ_ERR_ARG;
  if (nout) *nout = g->nout;
  if (f0) *f0 = g->F0;
  return QDEQ_OK;
}

// Copies between the API layout (unpadded tensors back to back) and the
// device layout (64-aligned tensors, W4 padded to noutp rows).
static int params_xfer(qdeq_guess *g, float *host, const float *dev_src,
                       float *dev_dst, bool to_host) {
  long long h = 0;
  for (int t = 0; t < 5; t++) {
    int rows, cols, rp;
    tensor_shape(g, t, &rows, &cols, &rp);
    const size_t n = (size_t)rows * cols;
    if (to_host)
      GHIP(g, hipMemcpy(host + h, dev_src + g->off[t], n * sizeof(float),
                        hipMemcpyDeviceToHost));
    else
      GHIP(g, hipMemcpy(dev_dst + g->off[t], host + h, n * sizeof(float),
                        hipMemcpyHostToDevice));
    h += (long long)n;
  }
  return QDEQ_OK;
}

int qdeq_guess_get_params(qdeq_guess *g, float *out) {
  if (!g || !out) return QDEQ_ERR_ARG;
  GHIP(g, hipStreamSynchronize(g->stream));
  return params_xfer(g, out, g->w, nullptr, true);
}

int qdeq_guess_set_params(qdeq_guess *g, const float *in) {
  if (!g || !in) return QDEQ_ERR_ARG;
  GHIP(g, hipStreamSynchronize(g->stream));
  int st = params_xfer(g, const_cast<float *>(in), nullptr, g->w, false);
  if (st) return st;
  refresh_bf16(g, g->stream);
  GHIP(g, hipMemsetAsync(g->m, 0, (size_t)g->ntot * sizeof(float), g->stream));
  GHIP(g, hipMemsetAsync(g->v, 0, (size_t)g->ntot * sizeof(float), g->stream));
  g->adam_t = 0;
  GHIP(g, hipStreamSynchronize(g->stream));
  return QDEQ_OK;
}

int qdeq_guess_pairs(const qdeq_guess *g, int *n) {
  if (!g || !n) return QDEQ_ERR_ARG;
  *n = g->count;
  return QDEQ_OK;
}

int qdeq_guess_clear_pairs(qdeq_guess *g) {
  if (!g) return QDEQ_ERR_ARG;
  g->count = 0;
  return QDEQ_OK;
}

int qdeq_guess_read_pairs(qdeq_guess *g, int first, int n, float *x, float *t,
                          double *scale) {
  if (!g || first < 0 || n < 0 || first + n > g->count) return QDEQ_ERR_ARG;
  GHIP(g, hipDeviceSynchronize());
  if (x)
    GHIP(g, hipMemcpy(x, g->xs + (size_t)first * g->nk,
                      (size_t)n * g->nk * sizeof(float), hipMemcpyDeviceToHost));
  if (t)
    GHIP(g, hipMemcpy2D(t, (size_t)g->nout * sizeof(float),
                        g->ts + (size_t)first * g->noutp,
                        (size_t)g->noutp * sizeof(float),
                        (size_t)g->nout * sizeof(float), (size_t)n,
                        hipMemcpyDeviceToHost));
  if (scale)
    GHIP(g, hipMemcpy(scale, g->scales + first, (size_t)n * sizeof(double),
                      hipMemcpyDeviceToHost));
  return QDEQ_OK;
}

int qdeq_guess_forward_backward(qdeq_guess *g, int batch, const float *x,
                                const float *t, double *loss, float *grad) {
  if (!g || batch < 1 || !x || !t) return QDEQ_ERR_ARG;
  const int Mb = (int)round_up(batch, 128);
  int st;
  if ((st = ensure_train_bufs(g, Mb))) return st;
  if ((st = upload_rows(g, g->X, Mb, g->nk, x, batch, g->nk))) return st;
  if ((st = upload_rows(g, g->Tg, Mb, g->noutp, t, batch, g->nout))) return st;
  if ((st = forward(g, g->stream, Mb))) return st;
  nn::mse(g->stream, g->Y, g->Tg, Mb, batch, g->nout, g->noutp, g->mse_part, g->loss);
  if ((st = backward(g, g->stream, Mb))) return st;
  GHIP(g, hipGetLastError());
  GHIP(g, hipStreamSynchronize(g->stream));
  if (loss) GHIP(g, hipMemcpy(loss, g->loss, sizeof(double), hipMemcpyDeviceToHost));
  if (grad) return params_xfer(g, grad, g->g, nullptr, true);
  return QDEQ_OK;
}

int qdeq_guess_forward(qdeq_guess *g, int batch, const float *x, float *y) {
  if (!g || batch < 1 || !x || !y) return QDEQ_ERR_ARG;
  const int Mb = (int)round_up(batch, 128);
  int st;
  if ((st = ensure_train_bufs(g, Mb))) return st;
  if ((st = upload_rows(g, g->X, Mb, g->nk, x, batch, g->nk))) return st;
  if ((st = forward(g, g->stream, Mb))) return st;
  GHIP(g, hipGetLastError());
  GHIP(g, hipMemcpy2DAsync(y, (size_t)g->nout * sizeof(float), g->Y,
                           (size_t)g->noutp * sizeof(float),
                           (size_t)g->nout * sizeof(float), (size_t)batch,
                           hipMemcpyDeviceToHost, g->stream));
  GHIP(g, hipStreamSynchronize(g->stream));
  return QDEQ_OK;
}

int qdeq_guess_train(qdeq_guess *g, const qdeq_train_config *tc,
                     double *final_loss) {
  if (!g || !tc || tc->epochs < 0 || tc->batch < 1 || !(tc->lr > 0) ||
      !(tc->lr_final > 0))
    return QDEQ_ERR_ARG;
  GHIP(g, hipDeviceSynchronize());
  std::vector<double> sc((size_t)g->count), mg((size_t)g->count);
  if (g->count) {
    GHIP(g, hipMemcpy(sc.data(), g->scales, sc.size() * sizeof(double),
                      hipMemcpyDeviceToHost));
    GHIP(g, hipMemcpy(mg.data(), g->mags, mg.size() * sizeof(double),
                      hipMemcpyDeviceToHost));
  }
  // A solve is accurate to its tolerance relative to the whole boundary
  // data, so an increment far below that data is dominated by solver error
  // in its target: such pairs are skipped.
  std::vector<int> valid;
  for (int i = 0; i < g->count; i++) {
    const double s = sc[(size_t)i];
    if (s > 0.0 && std::isfinite(s) && s >= kMinRelIncrement * mg[(size_t)i])
      valid.push_back(i);
  }
  if (valid.empty()) return gset_error(g, QDEQ_ERR_STATE, "no training pairs");
  const int B = std::min(tc->batch, (int)valid.size());
  const int Mb = (int)round_up(B, 128);
  int st;
  if ((st = ensure_train_bufs(g, Mb))) return st;
  if (g->idx_n < (int)valid.size()) {
    (void)hipFree(g->idx);
    g->idx = nullptr;
    GHIP(g, dalloc(&g->idx, valid.size()));
    g->idx_n = (int)valid.size();
  }
  const int nb = (int)(valid.size() / B);  // whole batches per epoch
  const long long total = (long long)nb * tc->epochs;
  std::mt19937_64 rng(tc->seed);
  double last = 0.0;
  long long step = 0;
  std::vector<double> losses;
  for (int e = 0; e < tc->epochs; e++) {
    std::shuffle(valid.begin(), valid.end(), rng);
    GHIP(g, hipMemcpyAsync(g->idx, valid.data(), valid.size() * sizeof(int),
                           hipMemcpyHostToDevice, g->stream));
    // the next epoch reshuffles the host list: this copy must be done first
    GHIP(g, hipStreamSynchronize(g->stream));
    const int nbl = std::min(nb, (1 << 16));
    for (int b = 0; b < nbl; b++, step++) {
      nn::gather(g->stream, g->xs, g->ts, g->idx + (size_t)b * B, B, Mb, g->nk,
                 g->noutp, g->X, g->Tg);
      if ((st = forward(g, g->stream, Mb))) return st;
      nn::mse(g->stream, g->Y, g->Tg, Mb, B, g->nout, g->noutp, g->mse_part,
              g->loss + b);
      if ((st = backward(g, g->stream, Mb))) return st;
      g->adam_t++;
      const double frac = total > 1 ? (double)step / (double)(total - 1) : 0.0;
      const double lr = tc->lr * std::pow(tc->lr_final / tc->lr, frac);
      const double c1 = 1.0 - std::pow(tc->beta1, (double)g->adam_t);
      const double c2 = 1.0 - std::pow(tc->beta2, (double)g->adam_t);
      nn::adam(g->stream, g->w, g->m, g->v, g->g, g->ntot, (float)lr,
               (float)tc->beta1, (float)tc->beta2, (float)tc->eps, (float)c1,
               (float)c2);
      refresh_bf16(g, g->stream);
    }
    GHIP(g, hipGetLastError());
    losses.assign((size_t)nbl, 0.0);
    GHIP(g, hipMemcpyAsync(losses.data(), g->loss, (size_t)nbl * sizeof(double),
                           hipMemcpyDeviceToHost, g->stream));
    GHIP(g, hipStreamSynchronize(g->stream));
    last = std::accumulate(losses.begin(), losses.end(), 0.0) / std::max(nbl, 1);
  }
  if (final_loss) *final_loss = last;
  return QDEQ_OK;
}

int qdeq_attach_guess(qdeq_ctx *c, qdeq_guess *g, int mode) {
  if (!c) return QDEQ_ERR_ARG;
  if (mode == QDEQ_GUESS_OFF || !g) {
    c->guess_mode = QDEQ_GUESS_OFF;
    guess_reset(c);
    return QDEQ_OK;
  }
  if (mode != QDEQ_GUESS_HARVEST && mode != QDEQ_GUESS_PREDICT)
    return set_error(c, QDEQ_ERR_ARG, "bad guess mode");
  if (c->proj_slots > 0)
    return set_error(c, QDEQ_ERR_STATE,
                     "learned guess and projection start are exclusive");
  if (g->cfg.device != c->cfg.device || g->nk != c->cfg.nk ||
      c->cfg.ni != c->cfg.nk || c->lv[0].N != g->nk || c->lv[0].K != g->nk)
    return set_error(c, QDEQ_ERR_ARG,
                     "guess network does not match the context grid (square "
                     "grids with ni == nk == the network's nk)");
  for (int l = 0; l < 4; l++) {
    const int in = l == 0 ? g->F0 : g->H, out = l == 3 ? g->noutp : g->H;
    if (!nn::gemm_shape_ok(nn::GEMM_CFG_SKINNY, 16, out, in, in, in))
      return set_error(c, QDEQ_ERR_ARG, "guess layer shape not covered by the tiles");
  }
  if (c->guess != g) {
    guess_free_ctx(c);
    const LevelDims &d = c->lv[0];
    const size_t n0 = (size_t)d.N * d.P;
    c->guess = g;
    QDEQ_HIP(c, dalloc(&c->wr, n0));
    QDEQ_HIP(c, dalloc(&c->gx, (size_t)16 * g->nk));
    QDEQ_HIP(c, dalloc(&c->gz, (size_t)16 * g->H));
    QDEQ_HIP(c, dalloc(&c->gy, (size_t)16 * g->noutp));
    c->gsplit_n = (size_t)64 * 16 * std::max(g->F0, g->H);
    QDEQ_HIP(c, dalloc(&c->gsplit, c->gsplit_n));
    QDEQ_HIP(c, dalloc(&c->gprev, (size_t)g->nk + 1));
    QDEQ_HIP(c, dalloc(&c->gbefore, (size_t)g->nout));
    QDEQ_HIP(c, dalloc(&c->gs, (size_t)4));
    for (int l = 0; l < 4; l++)
      QDEQ_HIP(c, dalloc(&c->gh[l], (size_t)16 * (l == 0 ? g->F0 : g->H)));
    const double one = 1.0;
    QDEQ_HIP(c, hipMemcpy(c->gs + 2, &one, sizeof(double), hipMemcpyHostToDevice));
    int st = solve_unit_remote(c, c->wr);
    if (st != QDEQ_OK) {
      guess_free_ctx(c);
      return st;
    }
  }
  c->guess_mode = mode;
  guess_reset(c);
  return QDEQ_OK;
}

int qdeq_guess_stats(qdeq_ctx *c, long long *predicted, long long *accepted,
                     long long *harvested) {
  if (!c) return QDEQ_ERR_ARG;
  if (predicted) *predicted = c->g_predicted;
  if (accepted) *accepted = c->g_accepted;
  if (harvested) *harvested = c->g_harvested;
  return QDEQ_OK;
}

int qdeq_guess_inject(qdeq_ctx *c, double factor) {
  if (!c || !c->guess) return set_error(c, QDEQ_ERR_STATE, "no guess attached");
  QDEQ_HIP(c, hipMemcpyAsync(c->gs + 2, &factor, sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
  QDEQ_HIP(c, hipStreamSynchronize(c->stream));
  return QDEQ_OK;
}

int qdeq_guess_infer(qdeq_ctx *c, const float *x, float *y) {
  if (!c || !x || !y) return set_error(c, QDEQ_ERR_ARG, "null argument");
  qdeq_guess *g = c->guess;
  if (!g) return set_error(c, QDEQ_ERR_STATE, "no guess attached");
  QDEQ_HIP(c, hipMemcpyAsync(c->gx, x, (size_t)g->nk * sizeof(float),
                             hipMemcpyHostToDevice, c->stream));
  nn::conv_forward(c->stream, c->gx, 16, g->nk, g->w + g->off[0], g->C, g->T,
                   g->f, g->P0, c->gh[0], nullptr);
  for (int l = 0; l < 3; l++) {
    const int in = l == 0 ? g->F0 : g->H;
    nn::gemm_nt(c->stream, nn::GEMM_CFG_SKINNY, c->gh[l], in, g->W[l], in, c->gz,
                g->H, 16, g->H, in, c->gsplit, c->gsplit_n, 64);
    nn::epilogue(c->stream, c->gz, 16, g->H, nn::EPI_RELU, nullptr, c->gh[l + 1],
                 nullptr, nullptr);
  }
  nn::gemm_nt(c->stream, nn::GEMM_CFG_SKINNY, c->gh[3], g->H, g->W[3], g->H, c->gy,
              g->noutp, 16, g->noutp, g->H, c->gsplit, c->gsplit_n, 64);
  QDEQ_HIP(c, hipGetLastError());
  QDEQ_HIP(c, hipMemcpyAsync(y, c->gy, (size_t)g->nout * sizeof(float),
                             hipMemcpyDeviceToHost, c->stream));
  QDEQ_HIP(c, hipStreamSynchronize(c->stream));
  return QDEQ_OK;
}

}  // extern "C"
