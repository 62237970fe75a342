// proj.hip -- projection start of the RK43 stage solves: the least-squares
// combination of the last few solutions whose fault data best match the new
// ones.  No reference counterpart: the reference iterates each solve from U
// as the previous solve left it, with the stage's boundary data written in
// (odeSolver.cpp:652-720, 1150-1188); that start remains the fallback here,
// and the combination is used only when its residual norm is finite and
// below the fallback's.
//
// The elliptic problem has f = 0, so a solution is linear in its boundary
// data: U = S(g, r) with g the fault column and r the remote value (uniform
// along the remote column).  With W_R = S(0, 1), Z = U - r W_R = S(g, 0).
// For new data (g*, r*) the start on the interior columns is
//     Z_0 + sum_i c_i (Z_i - Z_0) + r* W_R,
// Z_0 the newest stored Z, where c minimises
//     || (g* - g_0) - sum_i c_i (g_i - g_0) ||_2
// over the stored fault columns (classical Gram-Schmidt applied twice, in
// one workgroup, every sum in a fixed order).  The start's residual is then
// the response to the part of g* - g_0 outside the stored span, plus the
// stored solutions' own residuals.

#include <cmath>
#include <string>

#include "qdeq.h"
#include "qdeq_internal.h"

namespace qdeq {

constexpr int PROJ_THREADS = 256;
constexpr int PROJ_MAX_SLOTS = 8;
// a difference column whose norm after orthogonalisation falls below this
// fraction of its own norm is left out of the fit
constexpr double kProjDrop = 1e-10;

struct ProjSlots {
  int s[PROJ_MAX_SLOTS];  // ring slots, newest first
  int cnt;
};

namespace {

// Sums v[0..nv) over the 256 threads, the same result in every thread:
// within a wavefront a butterfly (lane l adds lane l^m for m = 32..1; the
// partners compute a + b and b + a, which are equal), then the four
// wavefront sums in wavefront order.  Loops run to the static bound so v
// stays in registers.
constexpr int PROJ_NV = PROJ_MAX_SLOTS + 1;

// v[0..nq) and, with last, v[PROJ_MAX_SLOTS] are summed.
__device__ inline void block_sums(double (&v)[PROJ_NV], int nq, bool last,
                                  double *sh) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < PROJ_NV; q++) {
    if (q < nq || (last && q == PROJ_MAX_SLOTS)) {
      double x = v[q];
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1) x = x + __shfl_xor(x, m, 64);
      v[q] = x;
      if (lane == 0) sh[w * PROJ_NV + q] = x;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PROJ_NV; q++)
    if (q < nq || (last && q == PROJ_MAX_SLOTS))
      v[q] = ((sh[q] + sh[PROJ_NV + q]) + sh[2 * PROJ_NV + q]) + sh[3 * PROJ_NV + q];
  __syncthreads();
}

// Orthogonalises x against Q columns 1..nq of lds twice (classical
// Gram-Schmidt, then again on the remainder); with want_norm the first pass
// also sums x.x.  r receives the two passes' coefficients added.
__device__ inline void cgs2(double *lds, int N, int nq, double *x, double *sh,
                            double (&r)[PROJ_NV], bool want_norm, double *nrm0) {
  const int tid = threadIdx.x;
  for (int pass = 0; pass < 2; pass++) {
    double v[PROJ_NV];
#pragma unroll
    for (int q = 0; q < PROJ_NV; q++) v[q] = 0.0;
    const bool nn = want_norm && pass == 0;
    for (int j = tid; j < N; j += PROJ_THREADS) {
      const double xj = x[j];
#pragma unroll
      for (int q = 0; q < PROJ_MAX_SLOTS; q++)
        if (q < nq) v[q] = v[q] + lds[(size_t)(q + 1) * N + j] * xj;
      if (nn) v[PROJ_MAX_SLOTS] = v[PROJ_MAX_SLOTS] + xj * xj;
    }
    block_sums(v, nq, nn, sh);
    if (nn) *nrm0 = sqrt(v[PROJ_MAX_SLOTS]);
    for (int j = tid; j < N; j += PROJ_THREADS) {
      double xj = x[j];
#pragma unroll
      for (int q = 0; q < PROJ_MAX_SLOTS; q++)
        if (q < nq) xj = xj - v[q] * lds[(size_t)(q + 1) * N + j];
      x[j] = xj;
    }
#pragma unroll
    for (int q = 0; q < PROJ_MAX_SLOTS; q++)
      if (q < nq) r[q] = pass == 0 ? v[q] : r[q] + v[q];
  }
}

__device__ inline double norm2(const double *x, int N, double *sh) {
  double v[PROJ_NV];
#pragma unroll
  for (int q = 0; q < PROJ_NV; q++) v[q] = 0.0;
  for (int j = threadIdx.x; j < N; j += PROJ_THREADS) v[0] = v[0] + x[j] * x[j];
  block_sums(v, 1, false, sh);
  return sqrt(v[0]);
}

// Coefficients of the start.  LDS: cnt columns of N doubles; column 0 holds
// d = g* - g_0, column i >= 1 holds g_i - g_0 on entry, and the kept
// orthonormal columns are packed into columns 1..nq as they are found (the
// a-th kept column is written over column a + 1 <= i).  Each thread owns
// the entries j = tid + 256 e of every column, so only the sums need the
// workgroup.  Output coef[i] = c_i for the stored solution i = 1..cnt-1 (0
// for a column left out), coef[PROJ_MAX_SLOTS] = ||d - fit|| / ||d||,
// coef[PROJ_MAX_SLOTS + 1] = columns kept.
__global__ __launch_bounds__(PROJ_THREADS) void k_proj_coef(
    const double *__restrict__ u, int N, int P, const double *__restrict__ g,
    ProjSlots sl, double *__restrict__ coef) {
  extern __shared__ __align__(16) double lds[];
  __shared__ double sh[4 * PROJ_NV];
  __shared__ double sR[PROJ_MAX_SLOTS * PROJ_MAX_SLOTS];
  __shared__ double sb[PROJ_MAX_SLOTS], sc[PROJ_MAX_SLOTS];
  __shared__ int skept[PROJ_MAX_SLOTS];
  const int tid = threadIdx.x;
  const int cnt = sl.cnt;
  const double *g0 = g + (size_t)sl.s[0] * N;
  for (int j = tid; j < N; j += PROJ_THREADS) {
    const double b0 = g0[j];
    lds[j] = u[(size_t)j * P] - b0;
    for (int i = 1; i < cnt; i++) lds[(size_t)i * N + j] = g[(size_t)sl.s[i] * N + j] - b0;
  }
  int nq = 0;
  double r[PROJ_NV];
  for (int i = 1; i < cnt; i++) {
    double *x = lds + (size_t)i * N;
    double nrm0 = 0.0;
    cgs2(lds, N, nq, x, sh, r, true, &nrm0);
    const double nrm = norm2(x, N, sh);
    if (nrm > 0.0 && nrm > kProjDrop * nrm0) {
      double *q = lds + (size_t)(nq + 1) * N;
      for (int j = tid; j < N; j += PROJ_THREADS) q[j] = x[j] / nrm;
      if (tid == 0) {
#pragma unroll
        for (int a = 0; a < PROJ_MAX_SLOTS; a++)
          if (a < nq) sR[a * PROJ_MAX_SLOTS + nq] = r[a];
        sR[nq * PROJ_MAX_SLOTS + nq] = nrm;
        skept[nq] = i;
      }
      nq++;
    }
  }
  // right-hand side b = Q^T d (twice, d updated in between)
  double nd0 = 0.0;
  cgs2(lds, N, nq, lds, sh, r, true, &nd0);
  const double nd = norm2(lds, N, sh);
  if (tid == 0) {
#pragma unroll
    for (int a = 0; a < PROJ_MAX_SLOTS; a++)
      if (a < nq) sb[a] = r[a];
    // back substitution R c = b over the kept columns
    for (int a = nq - 1; a >= 0; a--) {
      double acc = sb[a];
      for (int b = a + 1; b < nq; b++) acc = acc - sR[a * PROJ_MAX_SLOTS + b] * sc[b];
      sc[a] = acc / sR[a * PROJ_MAX_SLOTS + a];
    }
    for (int i = 0; i < PROJ_MAX_SLOTS; i++) coef[i] = 0.0;
    for (int a = 0; a < nq; a++) coef[skept[a]] = sc[a];
    coef[PROJ_MAX_SLOTS] = nd0 > 0.0 ? nd / nd0 : 0.0;
    coef[PROJ_MAX_SLOTS + 1] = (double)nq;
  }
}

// Start on the interior columns, the current U on the fault and remote
// columns (and the pitch padding).
__global__ void k_proj_apply(const double *__restrict__ u,
                             const double *__restrict__ wr,
                             const double *__restrict__ z, size_t zstride,
                             ProjSlots sl, const double *__restrict__ coef, int N,
                             int K, int P, double *__restrict__ cand) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (k >= P || j >= N) return;
  const size_t c = (size_t)j * P + k;
  if (k == 0 || k >= K - 1) {
    cand[c] = u[c];
    return;
  }
  const double z0 = z[(size_t)sl.s[0] * zstride + c];
  double acc = z0;
  for (int i = 1; i < sl.cnt; i++)
    acc = acc + coef[i] * (z[(size_t)sl.s[i] * zstride + c] - z0);
  cand[c] = acc + u[K - 1] * wr[c];
}

// Stores the solution just found: Z = U - r W_R and its fault column.
__global__ void k_proj_commit(const double *__restrict__ u,
                              const double *__restrict__ wr, int N, int K, int P,
                              double *__restrict__ z, double *__restrict__ g) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (k >= P || j >= N) return;
  const size_t c = (size_t)j * P + k;
  z[c] = u[c] - u[K - 1] * wr[c];
  if (k == 0) g[j] = u[c];
}

}  // namespace

static size_t proj_lds_bytes(int slots, int N) { return (size_t)slots * N * sizeof(double); }

void proj_reset(qdeq_ctx *c) {
  c->proj_count = 0;
  c->proj_head = 0;
}

void proj_free(qdeq_ctx *c) {
  void *p[] = {c->proj_z, c->proj_g, c->proj_coef, c->proj_wr};
  for (void *q : p) (void)hipFree(q);
  c->proj_z = c->proj_g = c->proj_coef = c->proj_wr = nullptr;
  c->proj_slots = 0;
  proj_reset(c);
}

static ProjSlots proj_order(const qdeq_ctx *c) {
  ProjSlots sl{};
  sl.cnt = c->proj_count;
  for (int i = 0; i < sl.cnt; i++)
    sl.s[i] = ((c->proj_head - 1 - i) % c->proj_slots + c->proj_slots) % c->proj_slots;
  return sl;
}

int proj_before_solve(qdeq_ctx *c) {
  if (c->proj_slots < 1 || c->proj_count < 1) return QDEQ_OK;
  const LevelDims &d = c->lv[0];
  const ProjSlots sl = proj_order(c);
  mbox_arm(c, SC_NORM_PRED);
  if (sl.cnt >= 2)
    hipLaunchKernelGGL(k_proj_coef, dim3(1), dim3(PROJ_THREADS),
                       proj_lds_bytes(sl.cnt, d.N), c->stream, c->d[0].u, d.N, d.P,
                       c->proj_g, sl, c->proj_coef);
  const dim3 blk(256), grd((d.P + 255) / 256, d.N);
  hipLaunchKernelGGL(k_proj_apply, grd, blk, 0, c->stream, c->d[0].u, c->proj_wr,
                     c->proj_z, (size_t)d.N * d.P, sl, c->proj_coef, d.N, d.K, d.P,
                     c->d[0].r);
  launch_resnorm_of(c, c->d[0].r, 2 * 1024, SC_NORM_PRED);
  QDEQ_HIP(c, hipGetLastError());
  c->cand_pending = true;
  c->g_predicted++;
  return QDEQ_OK;
}

int proj_after_solve(qdeq_ctx *c) {
  if (c->proj_slots < 1) return QDEQ_OK;
  const LevelDims &d = c->lv[0];
  const int slot = c->proj_head;
  const dim3 blk(256), grd((d.P + 255) / 256, d.N);
  hipLaunchKernelGGL(k_proj_commit, grd, blk, 0, c->stream, c->d[0].u, c->proj_wr,
                     d.N, d.K, d.P, c->proj_z + (size_t)slot * d.N * d.P,
                     c->proj_g + (size_t)slot * d.N);
  QDEQ_HIP(c, hipGetLastError());
  c->proj_head = (c->proj_head + 1) % c->proj_slots;
  if (c->proj_count < c->proj_slots) c->proj_count++;
  return QDEQ_OK;
}

}  // namespace qdeq

using namespace qdeq;

extern "C" {

int qdeq_projection_enable(qdeq_ctx *c, int slots) {
  if (!c) return QDEQ_ERR_ARG;
  if (slots == 0) {
    proj_free(c);
    return QDEQ_OK;
  }
  if (slots < 1 || slots > PROJ_MAX_SLOTS)
    return set_error(c, QDEQ_ERR_ARG, "projection slots must be 0..8");
  if (c->guess && c->guess_mode != QDEQ_GUESS_OFF)
    return set_error(c, QDEQ_ERR_STATE,
                     "projection start and learned guess are exclusive");
  const LevelDims &d = c->lv[0];
  if (proj_lds_bytes(slots, d.N) > c->lds_max)
    return set_error(c, QDEQ_ERR_ARG,
                     "projection slots x fault nodes exceed the workgroup LDS");
  proj_free(c);
  const size_t n0 = (size_t)d.N * d.P;
  QDEQ_HIP(c, hipMalloc(&c->proj_z, (size_t)slots * n0 * sizeof(double)));
  QDEQ_HIP(c, hipMalloc(&c->proj_g, (size_t)slots * d.N * sizeof(double)));
  QDEQ_HIP(c, hipMalloc(&c->proj_coef, (PROJ_MAX_SLOTS + 2) * sizeof(double)));
  QDEQ_HIP(c, hipMalloc(&c->proj_wr, n0 * sizeof(double)));
  QDEQ_HIP(c, hipMemset(c->proj_coef, 0, (PROJ_MAX_SLOTS + 2) * sizeof(double)));
  if (proj_lds_bytes(slots, d.N) > 65536)
    QDEQ_HIP(c, hipFuncSetAttribute((const void *)k_proj_coef,
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)proj_lds_bytes(slots, d.N)));
  int st = solve_unit_remote(c, c->proj_wr);
  if (st != QDEQ_OK) {
    proj_free(c);
    return st;
  }
  c->proj_slots = slots;
  proj_reset(c);
  return QDEQ_OK;
}

int qdeq_projection_fit(qdeq_ctx *c, const double *g_new, const double *g_old,
                        int cnt, double *coef) {
  if (!c || !g_new || !g_old || !coef) return QDEQ_ERR_ARG;
  const LevelDims &d = c->lv[0];
  const int N = d.N;
  if (cnt < 2 || cnt > PROJ_MAX_SLOTS)
    return set_error(c, QDEQ_ERR_ARG, "projection fit: cnt must be 2..8");
  if (proj_lds_bytes(cnt, N) > c->lds_max)
    return set_error(c, QDEQ_ERR_ARG, "projection fit exceeds the workgroup LDS");
  double *u = nullptr, *g = nullptr, *cf = nullptr;
  hipError_t e = hipMalloc(&u, (size_t)N * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&g, (size_t)cnt * N * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&cf, (PROJ_MAX_SLOTS + 2) * sizeof(double));
  if (e == hipSuccess && proj_lds_bytes(cnt, N) > 65536)
    e = hipFuncSetAttribute((const void *)k_proj_coef,
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)proj_lds_bytes(cnt, N));
  if (e == hipSuccess)
    e = hipMemcpy(u, g_new, (size_t)N * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(g, g_old, (size_t)cnt * N * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    ProjSlots sl{};
    sl.cnt = cnt;
    for (int i = 0; i < cnt; i++) sl.s[i] = i;
    // the new column as a one-column field (pitch 1)
    hipLaunchKernelGGL(k_proj_coef, dim3(1), dim3(PROJ_THREADS),
                       proj_lds_bytes(cnt, N), c->stream, u, N, 1, g, sl, cf);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  double h[PROJ_MAX_SLOTS + 2];
  if (e == hipSuccess)
    e = hipMemcpy(h, cf, sizeof(h), hipMemcpyDeviceToHost);
  (void)hipFree(u);
  (void)hipFree(g);
  (void)hipFree(cf);
  if (e != hipSuccess)
    return set_error(c, QDEQ_ERR_HIP, std::string("projection fit: ") + hipGetErrorString(e));
  for (int i = 0; i < cnt; i++) coef[i] = h[i];
  coef[cnt] = h[PROJ_MAX_SLOTS];
  coef[cnt + 1] = h[PROJ_MAX_SLOTS + 1];
  return QDEQ_OK;
}

int qdeq_projection_stats(qdeq_ctx *c, long long *predicted, long long *accepted,
                          double *fit_ratio, int *columns) {
  if (!c) return QDEQ_ERR_ARG;
  if (predicted) *predicted = c->g_predicted;
  if (accepted) *accepted = c->g_accepted;
  if (fit_ratio || columns) {
    double h[2] = {0.0, 0.0};
    if (c->proj_coef) {
      QDEQ_HIP(c, hipStreamSynchronize(c->stream));
      QDEQ_HIP(c, hipMemcpy(h, c->proj_coef + PROJ_MAX_SLOTS, sizeof(h),
                            hipMemcpyDeviceToHost));
    }
    if (fit_ratio) *fit_ratio = h[0];
    if (columns) *columns = (int)h[1];
  }
  return QDEQ_OK;
}

}  // extern "C"
