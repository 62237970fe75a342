// proj.hip -- projection start of the RK43 stage solves: the least-squares
// combination of the last few solutions whose fault data best match the new
// ones.  No reference counterpart: the reference iterates each solve from U
// as the previous solve left it, with the stage's boundary data written in
// (odeSolver.cpp:652-720, 1150-1188); that start remains the fallback here,
// and the combination is used only when its residual norm is finite and
// below the fallback's.
//
// The elliptic problem has f = 0, so a solution is linear in its boundary
// data: U = S(g, r) with g the fault column and r the remote value (the
// remote column is r plus a fixed per-row shift).  With W_R = S(0, 1) the
// response to a unit change of r, the start for new data (g*, r*) on the
// interior columns is
//     U_0 + sum_i c_i (U_i - U_0) + w W_R,   w = r* - r_0 - sum_i c_i (r_i - r_0),
// U_0 the newest stored solution, where c minimises
//     || (g* - g_0) - sum_i c_i (g_i - g_0) ||_2
// over the stored fault columns (CholeskyQR applied twice, in one
// workgroup, every sum in a fixed order).  The start's residual is then
// the response to the part of g* - g_0 outside the stored span, plus the
// stored solutions' own residuals.  The ring keeps the solutions
// themselves, so the previous solution (the plain start's interior) is
// always at hand.

#include <cmath>
#include <string>

#include "qdeq.h"
#include "qdeq_internal.h"

namespace qdeq {

constexpr int PROJ_THREADS = 256;
constexpr int PROJ_MAX_SLOTS = 8;
struct ProjSlots {
  int s[PROJ_MAX_SLOTS];  // ring slots, newest first
  int cnt;
};

namespace {

// The fit is CholeskyQR applied twice to the augmented matrix
// A = [D_1 .. D_{m-1} | d] (columns in LDS, N entries each): G1 = A^T A by
// one pass over the entries and one tree sum, its Cholesky factor R1
// (G1 = R1^T R1), Q = A R1^{-1} in place, G2 = Q^T Q, R2 = chol(G2) and
// R = R2 R1; c solves R_DD c = R_Dd and ||d - D c|| = R_dd.  The small
// factorisations run in registers in every thread alike.  A D column whose
// part independent of the columns before it has a squared norm of at most
// kProjDrop2 of its own in the first factorisation (a norm ratio of 1e-6)
// is left out; the second factorisation restores orthogonality lost to the
// conditioning of the first.
//
// Order of every sum (restated on the CPU for bit-for-bit checks):
// * G_ab: thread t adds A_a[j] * A_b[j] for j = t, t+256, ... from 0.0; the
//   256 partial sums are then added pairwise in LDS, entry t += entry t+s
//   for s = 128, 64, ..., 1;
// * Cholesky, column by column: R_ki = (G_ki - sum_{l<k} R_lk R_li) * w_k
//   and p_i = G_ii - sum_{l<i} R_li R_li, terms subtracted in increasing l,
//   with R_ii = sqrt(p_i) and w_i = 1 / R_ii; a column left out has a zero
//   column of R and w = 0 (so its row of R is zero and every term that
//   involves it is zero);
// * Q_i = (A_i - sum_{k<i} R_ki Q_k) * w_i, terms subtracted in increasing k;
// * R_ij = sum_{k=i..j} R2_ik R1_kj from 0.0 in increasing k;
// * back substitution from the last D column to the first, each row
//   subtracting R_ik c_k in increasing k; c = 0 for a column left out.
constexpr double kProjDrop2 = 1e-12;

// Packed upper-triangle index of (a, b), a <= b < M.
template <int M>
__host__ __device__ constexpr int pidx(int a, int b) {
  return a * M - a * (a - 1) / 2 + (b - a);
}

// Gram entries of the M columns of A into Gs (LDS), then every thread reads
// them into G.  Pt: M(M+1)/2 * 256 doubles of LDS.
template <int M>
__device__ inline void gram(const double *A, int N, double *Pt, double *Gs,
                            double (&G)[M * (M + 1) / 2]) {
  constexpr int NP = M * (M + 1) / 2;
  const int tid = threadIdx.x;
  {
    double v[NP];
#pragma unroll
    for (int q = 0; q < NP; q++) v[q] = 0.0;
    for (int j = tid; j < N; j += PROJ_THREADS) {
      double x[M];
#pragma unroll
      for (int a = 0; a < M; a++) x[a] = A[(size_t)a * N + j];
#pragma unroll
      for (int a = 0; a < M; a++)
#pragma unroll
        for (int b = a; b < M; b++) v[pidx<M>(a, b)] = v[pidx<M>(a, b)] + x[a] * x[b];
    }
#pragma unroll
    for (int q = 0; q < NP; q++) Pt[q * PROJ_THREADS + tid] = v[q];
  }
  __syncthreads();
  for (int ls = 7; ls >= 0; ls--) {
    const int s = 1 << ls;
    for (int idx = tid; idx < NP * s; idx += PROJ_THREADS) {
      const int q = idx >> ls, t = idx & (s - 1);
      double *p = Pt + q * PROJ_THREADS;
      p[t] = p[t] + p[t + s];
    }
    __syncthreads();
  }
  if (tid < NP) Gs[tid] = Pt[tid * PROJ_THREADS];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NP; q++) G[q] = Gs[q];
  __syncthreads();
}

// Upper Cholesky factor (packed) of the M x M matrix G over the D columns
// in keep (bit i) and column M-1 (d, always factored, R_dd =
// sqrt(max(p, 0))).  With drop, D columns failing the rule are removed from
// keep and their column of R is zero.
template <int M>
__device__ inline void cholesky(const double (&G)[M * (M + 1) / 2], bool drop,
                                unsigned &keep, double (&R)[M * (M + 1) / 2],
                                double (&inv)[M]) {
#pragma unroll
  for (int q = 0; q < M * (M + 1) / 2; q++) R[q] = 0.0;
#pragma unroll
  for (int i = 0; i < M; i++) {
    inv[i] = 0.0;
    // a column left out has w = 0, so its row of R is zero and its terms
    // in the sums below are zero
#pragma unroll
    for (int k = 0; k < i; k++) {
      double s = G[pidx<M>(k, i)];
#pragma unroll
      for (int l = 0; l < k; l++) s = s - R[pidx<M>(l, k)] * R[pidx<M>(l, i)];
      R[pidx<M>(k, i)] = s * inv[k];
    }
    double p = G[pidx<M>(i, i)];
#pragma unroll
    for (int l = 0; l < i; l++) p = p - R[pidx<M>(l, i)] * R[pidx<M>(l, i)];
    const bool used = i == M - 1 || ((keep >> i) & 1u);
    const bool dropped =
        !used || (i < M - 1 && drop &&
                  !(G[pidx<M>(i, i)] > 0.0 && p > kProjDrop2 * G[pidx<M>(i, i)]));
    if (dropped) {
      keep &= ~(1u << i);
#pragma unroll
      for (int k = 0; k <= i; k++) R[pidx<M>(k, i)] = 0.0;
    } else {
      R[pidx<M>(i, i)] = p > 0.0 ? sqrt(p) : 0.0;
      inv[i] = p > 0.0 ? 1.0 / R[pidx<M>(i, i)] : 0.0;
    }
  }
}

// Coefficients of the start for M = cnt stored solutions.  LDS: M columns
// of N doubles (column i < M-1 holds g_{i+1} - g_0, column M-1 holds
// d = g* - g_0), then the tree sums.  Output coef[i] = c_i, the weight of
// (Z_i - Z_0), i = 1..cnt-1 (0 for a column left out, and beyond cnt),
// coef[PROJ_MAX_SLOTS] = ||d - fit|| / ||d||, coef[PROJ_MAX_SLOTS + 1] =
// columns kept.
template <int M>
__global__ __launch_bounds__(PROJ_THREADS) void k_proj_coef(
    const double *__restrict__ u, int N, int P, const double *__restrict__ g,
    ProjSlots sl, double *__restrict__ coef) {
  constexpr int NP = M * (M + 1) / 2;
  extern __shared__ __align__(16) double lds[];
  __shared__ double Gs[NP];
  const int tid = threadIdx.x;
  double *A = lds;
  double *Pt = lds + (size_t)M * N;
  const double *g0 = g + (size_t)sl.s[0] * N;
  for (int j = tid; j < N; j += PROJ_THREADS) {
    const double b0 = g0[j];
#pragma unroll
    for (int i = 0; i < M - 1; i++) A[(size_t)i * N + j] = g[(size_t)sl.s[i + 1] * N + j] - b0;
    A[(size_t)(M - 1) * N + j] = u[(size_t)j * P] - b0;
  }
  __syncthreads();
  unsigned keep = (1u << (M - 1)) - 1u;  // the D columns
  double G[NP], R1[NP], inv[M];
  gram<M>(A, N, Pt, Gs, G);
  const double dd = G[pidx<M>(M - 1, M - 1)];
  cholesky<M>(G, true, keep, R1, inv);
  // Q = A R1^{-1} on each thread's own entries
  for (int j = tid; j < N; j += PROJ_THREADS) {
    double x[M];
#pragma unroll
    for (int i = 0; i < M; i++) {
      double q = A[(size_t)i * N + j];
#pragma unroll
      for (int k = 0; k < i; k++) q = q - R1[pidx<M>(k, i)] * x[k];
      x[i] = q * inv[i];
      A[(size_t)i * N + j] = x[i];
    }
  }
  __syncthreads();
  double R2[NP];
  gram<M>(A, N, Pt, Gs, G);
  cholesky<M>(G, false, keep, R2, inv);
  // R = R2 R1 (into G), then the back substitution
#pragma unroll
  for (int i = 0; i < M; i++)
#pragma unroll
    for (int j2 = i; j2 < M; j2++) {
      double s = 0.0;
#pragma unroll
      for (int k = i; k <= j2; k++) s = s + R2[pidx<M>(i, k)] * R1[pidx<M>(k, j2)];
      G[pidx<M>(i, j2)] = s;
    }
  double c[M];
#pragma unroll
  for (int i = M - 2; i >= 0; i--) {
    double acc = G[pidx<M>(i, M - 1)];
#pragma unroll
    for (int k = i + 1; k < M - 1; k++) acc = acc - G[pidx<M>(i, k)] * c[k];
    c[i] = ((keep >> i) & 1u) ? acc / G[pidx<M>(i, i)] : 0.0;
  }
  const int kept = __builtin_popcount(keep);
  if (tid == 0) {
    coef[0] = 0.0;
#pragma unroll
    for (int i = 1; i < PROJ_MAX_SLOTS; i++) coef[i] = i < M ? c[i - 1] : 0.0;
    const double rd = G[pidx<M>(M - 1, M - 1)];
    coef[PROJ_MAX_SLOTS] = dd > 0.0 ? rd / sqrt(dd) : 0.0;
    coef[PROJ_MAX_SLOTS + 1] = (double)kept;
  }
}

// Start on the interior columns (1..K-2) into cand; with copy_bnd the fault
// and remote columns (and the pitch padding) are copied from u as well.
// cand may be u itself (interior points only are written then).  The weight
// w of W_R is formed by every thread alike, in the order above.
__global__ void k_proj_apply(const double *u, const double *__restrict__ wr,
                             const double *__restrict__ ring, size_t rstride,
                             const double *__restrict__ rr, ProjSlots sl,
                             const double *__restrict__ coef, int N, int K, int P,
                             double *cand, int copy_bnd, double inject) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (k >= P || j >= N) return;
  const size_t c = (size_t)j * P + k;
  if (k == 0 || k >= K - 1) {
    if (copy_bnd) cand[c] = u[c];
    return;
  }
  const double r0 = rr[sl.s[0]];
  double w = u[K - 1] - r0;
  for (int i = 1; i < sl.cnt; i++) w = w - coef[i] * (rr[sl.s[i]] - r0);
  const double u0 = ring[(size_t)sl.s[0] * rstride + c];
  double acc = u0;
  for (int i = 1; i < sl.cnt; i++)
    acc = acc + coef[i] * (ring[(size_t)sl.s[i] * rstride + c] - u0);
  const double v = acc + w * wr[c];
  // fault injection for tests (qdeq_projection_inject): the start's change
  // from the newest solution scaled
  cand[c] = inject == 1.0 ? v : u0 + inject * (v - u0);
}

// Stores the solution just found, its fault column and its remote value.
__global__ void k_proj_commit(const double *__restrict__ u, int N, int K, int P,
                              double *__restrict__ dst, double *__restrict__ g,
                              double *__restrict__ r) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (k >= P || j >= N) return;
  const size_t c = (size_t)j * P + k;
  dst[c] = u[c];
  if (k == 0) g[j] = u[c];
  if (j == 0 && k == K - 1) r[0] = u[c];
}

}  // namespace

static size_t proj_lds_bytes(int slots, int N) {
  return ((size_t)slots * N + (size_t)slots * (slots + 1) / 2 * PROJ_THREADS) *
         sizeof(double);
}

using CoefKernel = void (*)(const double *, int, int, const double *, ProjSlots, double *);

static CoefKernel coef_kernel(int cnt) {
  switch (cnt) {
    case 2: return k_proj_coef<2>;
    case 3: return k_proj_coef<3>;
    case 4: return k_proj_coef<4>;
    case 5: return k_proj_coef<5>;
    case 6: return k_proj_coef<6>;
    case 7: return k_proj_coef<7>;
    default: return k_proj_coef<8>;
  }
}

// Opts every fit kernel up to cnt columns into the LDS it needs.
static hipError_t coef_lds_attr(int cnt, int N) {
  for (int k = 2; k <= cnt; k++) {
    const size_t b = proj_lds_bytes(k, N);
    if (b <= 65536) continue;
    const hipError_t e = hipFuncSetAttribute((const void *)coef_kernel(k),
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)b);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

void proj_reset(qdeq_ctx *c) {
  c->proj_count = 0;
  c->proj_head = 0;
}

void proj_free(qdeq_ctx *c) {
  void *p[] = {c->proj_u, c->proj_g, c->proj_r, c->proj_coef, c->proj_wr};
  for (void *q : p) (void)hipFree(q);
  c->proj_u = c->proj_g = c->proj_r = c->proj_coef = c->proj_wr = nullptr;
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
  // the start's norms come with the solve (the same test as mg_begin's)
  const bool dev = device_start_cand(c);
  if (!dev) mbox_arm(c, SC_NORM_PRED);
  if (sl.cnt >= 2)
    hipLaunchKernelGGL(coef_kernel(sl.cnt), dim3(1), dim3(PROJ_THREADS),
                       proj_lds_bytes(sl.cnt, d.N), c->stream, c->d[0].u, d.N, d.P,
                       c->proj_g, sl, c->proj_coef);
  const dim3 blk(256), grd((d.P + 255) / 256, d.N);
  // on the device path the start is written over u's interior (the plain
  // start's interior is the newest stored solution); on the host path into r
  hipLaunchKernelGGL(k_proj_apply, grd, blk, 0, c->stream, c->d[0].u, c->proj_wr,
                     c->proj_u, (size_t)d.N * d.P, c->proj_r, sl, c->proj_coef,
                     d.N, d.K, d.P, dev ? c->d[0].u : c->d[0].r, dev ? 0 : 1,
                     c->proj_inject);
  if (!dev) launch_resnorm_of(c, c->d[0].r, 2 * kNormCap, SC_NORM_PRED);
  QDEQ_HIP(c, hipGetLastError());
  c->cand_pending = true;
  c->cand_in_u = dev;
  c->plain_interior = c->proj_u + (size_t)sl.s[0] * d.N * d.P;
  c->g_predicted++;
  return QDEQ_OK;
}

int proj_after_solve(qdeq_ctx *c) {
  if (c->proj_slots < 1) return QDEQ_OK;
  const LevelDims &d = c->lv[0];
  const int slot = c->proj_head;
  const dim3 blk(256), grd((d.P + 255) / 256, d.N);
  hipLaunchKernelGGL(k_proj_commit, grd, blk, 0, c->stream, c->d[0].u, d.N, d.K, d.P,
                     c->proj_u + (size_t)slot * d.N * d.P,
                     c->proj_g + (size_t)slot * d.N, c->proj_r + slot);
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
  QDEQ_HIP(c, hipMalloc(&c->proj_u, (size_t)slots * n0 * sizeof(double)));
  QDEQ_HIP(c, hipMalloc(&c->proj_g, (size_t)slots * d.N * sizeof(double)));
  QDEQ_HIP(c, hipMalloc(&c->proj_r, (size_t)slots * sizeof(double)));
  QDEQ_HIP(c, hipMalloc(&c->proj_coef, (PROJ_MAX_SLOTS + 2) * sizeof(double)));
  QDEQ_HIP(c, hipMalloc(&c->proj_wr, n0 * sizeof(double)));
  QDEQ_HIP(c, hipMemset(c->proj_coef, 0, (PROJ_MAX_SLOTS + 2) * sizeof(double)));
  QDEQ_HIP(c, coef_lds_attr(slots, d.N));
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
  if (e == hipSuccess) e = coef_lds_attr(cnt, N);
  if (e == hipSuccess)
    e = hipMemcpy(u, g_new, (size_t)N * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(g, g_old, (size_t)cnt * N * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    ProjSlots sl{};
    sl.cnt = cnt;
    for (int i = 0; i < cnt; i++) sl.s[i] = i;
    // the new column as a one-column field (pitch 1)
    hipLaunchKernelGGL(coef_kernel(cnt), dim3(1), dim3(PROJ_THREADS),
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

int qdeq_projection_inject(qdeq_ctx *c, double factor) {
  if (!c) return QDEQ_ERR_ARG;
  c->proj_inject = factor;
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
