// mg.hip -- geometric multigrid for div(mu grad u) = 0 on gfx950.
//
// Method (XuSun-SUSTech/QDEqCycle_FDM odeSolver.cpp:652-1147): V-cycle with
// red-black Gauss-Seidel smoothing, full-weighting restriction, bilinear
// prolongation, level-0 residual norm as the stop test.  The operator on
// each level is stored as compact face-coefficient arrays (cx: west face,
// cz: south face) extracted once from the doubled mu_equi array, so the
// sweep reads 2 coefficient words per point instead of 4 stride-2 gathers.
// Sums are formed in the reference's order and the build uses
// -ffp-contract=off, so the FP64 reference cycle reproduces the reference's
// iterates bit for bit.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <vector>

#include "qdeq_internal.h"

namespace qdeq {

constexpr int BX = 64, BY = 4;          // 256-thread tiles, 64-wide rows
// NORM_BLOCKS: capacity of the row-partial layout; a launch uses
// norm_blocks(c) = min(N, NORM_BLOCKS) blocks, one row each up to 2048 rows
// (one block with twice the rows of the others would set the time)
constexpr int NORM_BLOCKS = kNormCap, NORM_THREADS = 256;

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// One colour of RBGS (rbgs_kernel<phase>, odeSolver.cpp:757-861).  Threads
// are indexed over the points of one colour only: point k = 2t + ((j+c)&1),
// so every lane does work.  The update reads only the other colour.  Top and
// bottom rows mirror the vertical neighbour (the reference's Neumann rows);
// the coarse-level coefficient quirk of the reference (one vertical
// coefficient used twice) is encoded in cz's first and last rows.
template <typename T, bool HAS_F, bool EXACT>
__global__ __launch_bounds__(BX *BY) void k_rbgs(
    T *__restrict__ u, const T *__restrict__ f, const T *__restrict__ cx,
    const T *__restrict__ cz, int N, int K, int P, T DH2, T relax,
    int colour) {
  const int j = blockIdx.y * BY + threadIdx.y;
  const int t = blockIdx.x * BX + threadIdx.x;
  if (j >= N) return;
  const int k = 2 * t + ((j + colour) & 1);
  if (k < 1 || k > K - 2) return;
  const int jS = (j == 0) ? 1 : j - 1;
  const int jN = (j == N - 1) ? N - 2 : j + 1;
  const size_t c = (size_t)j * P + k;
  const T m1 = cx[c], m2 = cx[c + 1], m3 = cz[c], m4 = cz[c + P];
  const T uW = u[c - 1], uE = u[c + 1];
  const T uS = u[(size_t)jS * P + k], uN = u[(size_t)jN * P + k];
  T num = m1 * uW + m2 * uE + m3 * uS + m4 * uN;
  if (HAS_F) num = num - DH2 * f[c];
  if (EXACT) {
    u[c] = (relax / (m1 + m2 + m3 + m4)) * num + (T(1) - relax) * u[c];
  } else {
    const T unew = num / (m1 + m2 + m3 + m4);
    u[c] = (relax == T(1)) ? unew : u[c] + relax * (unew - u[c]);
  }
}

// r = f - A u on one level (residual_kernel, odeSolver.cpp:863-951).
template <typename T, bool HAS_F>
__global__ __launch_bounds__(BX *BY) void k_residual(
    T *__restrict__ r, const T *__restrict__ u, const T *__restrict__ f,
    const T *__restrict__ cx, const T *__restrict__ cz, int N, int K, int P,
    T DH2) {
  const int j = blockIdx.y * BY + threadIdx.y;
  const int k = blockIdx.x * BX + threadIdx.x + 1;
  if (j >= N || k > K - 2) return;
  const int jS = (j == 0) ? 1 : j - 1;
  const int jN = (j == N - 1) ? N - 2 : j + 1;
  const size_t c = (size_t)j * P + k;
  const T m1 = cx[c], m2 = cx[c + 1], m3 = cz[c], m4 = cz[c + P];
  const T fv = HAS_F ? f[c] : T(0);
  r[c] = fv - (m1 * u[c - 1] + m2 * u[c + 1] + m3 * u[(size_t)jS * P + k] +
               m4 * u[(size_t)jN * P + k] - (m1 + m2 + m3 + m4) * u[c]) /
                  DH2;
}

// Full weighting to the coarse level, zeroing the coarse correction
// (restriction_kernel, odeSolver.cpp:953-1015).  The first/last coarse rows
// use the single interior fine neighbour row twice.
template <typename T>
__global__ __launch_bounds__(BX *BY) void k_restrict(
    T *__restrict__ r2, T *__restrict__ e2, const T *__restrict__ r, int Nc,
    int Kc, int Pc, int Pf) {
  const int j = blockIdx.y * BY + threadIdx.y;
  const int k = blockIdx.x * BX + threadIdx.x + 1;
  if (j >= Nc || k > Kc - 2) return;
  const int jS = (j == 0) ? 1 : 2 * j - 1;
  const int jN = (j == Nc - 1) ? 2 * j - 1 : 2 * j + 1;
  const T *rc = r + (size_t)(2 * j) * Pf;
  const T *rs = r + (size_t)jS * Pf;
  const T *rn = r + (size_t)jN * Pf;
  const int kf = 2 * k;
  const T v = T(0.0625) *
              (T(4.0) * rc[kf] + T(2.0) * rc[kf - 1] + T(2.0) * rc[kf + 1] +
               T(2.0) * rs[kf] + T(2.0) * rn[kf] + rs[kf - 1] + rs[kf + 1] +
               rn[kf - 1] + rn[kf + 1]);
  r2[(size_t)j * Pc + k] = v;
  e2[(size_t)j * Pc + k] = T(0);
}

// u += P e (prolongate_kernel, odeSolver.cpp:1017-1043): one thread per
// coarse cell writes its 4 fine points.  The reference skips the last fine
// row (FULL_TOP = false); the fast cycle also corrects it.
template <typename T, bool FULL_TOP>
__global__ __launch_bounds__(BX *BY) void k_prolong(
    T *__restrict__ u, const T *__restrict__ e, int Nc, int Kc, int Pc,
    int Pf) {
  const int j = blockIdx.y * BY + threadIdx.y;
  const int k = blockIdx.x * BX + threadIdx.x;
  const int jmax = FULL_TOP ? Nc : Nc - 1;
  if (j >= jmax || k >= Kc - 1) return;
  const T *ec = e + (size_t)j * Pc + k;
  T *uf = u + (size_t)(2 * j) * Pf + 2 * k;
  const T e00 = ec[0], e01 = ec[1];
  uf[0] += e00;
  uf[1] += T(0.5) * (e00 + e01);
  if (!FULL_TOP || j < Nc - 1) {
    const T e10 = ec[Pc], e11 = ec[Pc + 1];
    T *uf1 = uf + Pf;
    uf1[0] += T(0.5) * (e00 + e10);
    uf1[1] += T(0.25) * (e00 + e01 + e10 + e11);
  }
}

// Workgroup sum (wavefront butterflies, wavefronts in order), every thread.
__device__ inline double block_sum_256(double v, double *sh) {
  return block_sum1(v, sh);
}

// Residual of the level-0 operator at interior point (j, k), f = 0.
__device__ inline double residual_at(const double *__restrict__ u,
                                     const double *__restrict__ cx,
                                     const double *__restrict__ cz, int j, int jS,
                                     int jN, int k, int P, double DH2) {
  const size_t c = (size_t)j * P + k;
  const double m1 = cx[c], m2 = cx[c + 1], m3 = cz[c], m4 = cz[c + P];
  return 0.0 - (m1 * u[c - 1] + m2 * u[c + 1] + m3 * u[(size_t)jS * P + k] +
                m4 * u[(size_t)jN * P + k] - (m1 + m2 + m3 + m4) * u[c]) /
                   DH2;
}

// The same residual with the interior values from ui and the fault and
// remote column values from ub (a start whose interior lives elsewhere).
__device__ inline double residual_at2(const double *__restrict__ ui,
                                      const double *__restrict__ ub,
                                      const double *__restrict__ cx,
                                      const double *__restrict__ cz, int j, int jS,
                                      int jN, int k, int K, int P, double DH2) {
  const size_t c = (size_t)j * P + k;
  const double m1 = cx[c], m2 = cx[c + 1], m3 = cz[c], m4 = cz[c + P];
  const double uW = k == 1 ? ub[c - 1] : ui[c - 1];
  const double uE = k == K - 2 ? ub[c + 1] : ui[c + 1];
  return 0.0 - (m1 * uW + m2 * uE + m3 * ui[(size_t)jS * P + k] +
                m4 * ui[(size_t)jN * P + k] - (m1 + m2 + m3 + m4) * ui[c]) /
                   DH2;
}

// Level-0 residual norm, pass 1: fixed-grid partial sums of r^2 (the
// reference uses calcuDelta_dev + cublasDnrm2, odeSolver.cpp:693-700).
// Optionally stores r as FP32 (the IR right-hand side) and zeroes the FP32
// correction.  Deterministic: fixed grid, fixed per-thread order, fixed tree.
template <bool STORE32>
__global__ __launch_bounds__(NORM_THREADS) void k_resnorm_partial(
    const double *__restrict__ u, const double *__restrict__ cx,
    const double *__restrict__ cz, int N, int K, int P, double DH2,
    double *__restrict__ partial, float *__restrict__ r32,
    float *__restrict__ e32, const int *__restrict__ gate) {
  if (gate && gate[0]) return;  // batched solve already converged
  __shared__ double sh[NORM_THREADS];
  double acc = 0.0;
  constexpr int UNR = 4;  // as in k_start_partials
  for (int j = blockIdx.x; j < N; j += gridDim.x) {
    const int jS = (j == 0) ? 1 : j - 1;
    const int jN = (j == N - 1) ? N - 2 : j + 1;
    for (int k0 = 1 + threadIdx.x; k0 <= K - 2; k0 += UNR * NORM_THREADS) {
      double r[UNR];
      // points past the row end evaluate the last interior point instead
      // (not added), so every load is issued before the first use
#pragma unroll
      for (int q = 0; q < UNR; q++) {
        const int k = min(k0 + q * NORM_THREADS, K - 2);
        r[q] = residual_at(u, cx, cz, j, jS, jN, k, P, DH2);
      }
#pragma unroll
      for (int q = 0; q < UNR; q++) {
        const int k = k0 + q * NORM_THREADS;
        if (k <= K - 2) {
          acc += r[q] * r[q];
          if (STORE32) {
            const size_t c = (size_t)j * P + k;
            r32[c] = (float)r[q];
            e32[c] = 0.0f;
          }
        }
      }
    }
  }
  const double s = block_sum_256(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Boundary-data residual ||r(u_interior = 0)||^2 partials: only the columns
// next to the Dirichlet columns are non-zero.  One thread per row.
__global__ __launch_bounds__(NORM_THREADS) void k_bdres_partial(
    const double *__restrict__ u, const double *__restrict__ cx, int N, int K,
    int P, double DH2, double *__restrict__ partial) {
  __shared__ double sh[NORM_THREADS];
  double acc = 0.0;
  for (int j = blockIdx.x * NORM_THREADS + threadIdx.x; j < N;
       j += gridDim.x * NORM_THREADS) {
    const size_t row = (size_t)j * P;
    for (int pass = 0; pass < 2; pass++) {
      const int k = pass == 0 ? 1 : K - 2;
      if (pass == 1 && K - 2 == 1) break;
      const double uW = (k - 1 == 0) ? u[row] : 0.0;
      const double uE = (k + 1 == K - 1) ? u[row + K - 1] : 0.0;
      const double m1 = cx[row + k], m2 = cx[row + k + 1];
      const double r = 0.0 - (m1 * uW + m2 * uE) / DH2;
      acc += r * r;
    }
  }
  const double s = block_sum_256(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Pass 2: one block sums the partials in a fixed order, writes sqrt.
// sqrt of the sum of n partials into out[0] and, when given, into the
// host-mapped mailbox slot mb[0] (read by the host without a stream sync).
__global__ __launch_bounds__(NORM_THREADS) void k_norm_final(
    const double *__restrict__ partial, int n, double *__restrict__ out,
    double *__restrict__ mb) {
  __shared__ double sh[NORM_THREADS];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += NORM_THREADS) acc += partial[i];
  const double s = block_sum_256(acc, sh);
  if (threadIdx.x == 0) {
    const double v = sqrt(s);
    out[0] = v;
    if (mb) mb[0] = v;
  }
}

// Stop test of one V-cycle of a batched solve.  gate[0]: stop (the solve
// converged or failed; the remaining cycles of the batch skip), gate[1]:
// V-cycles done, gate[2]: outcome (0 continue, 1 converged, 2 non-finite
// residual, 3 V-cycle cap), gate[3]: sequence number of the last cycle seen.
// The host waits for mb[MB_SEQ] to reach the batch's last sequence number;
// a skipped cycle only advances it.  The threshold is tol * the boundary-
// data norm (REL) or tol (ABS), the comparisons those of the host loop.
__global__ __launch_bounds__(NORM_THREADS) void k_norm_gate(
    const double *__restrict__ partial, int n, double *__restrict__ scal,
    double *__restrict__ mb, int *__restrict__ gate, int rel, double tol,
    int cap) {
  __shared__ double sh[NORM_THREADS];
  const bool skip = gate[0] != 0;
  double acc = 0.0;
  if (!skip)
    for (int i = threadIdx.x; i < n; i += NORM_THREADS) acc += partial[i];
  const double s = block_sum_256(acc, sh);
  if (threadIdx.x == 0) {
    const int seq = gate[3] + 1;
    gate[3] = seq;
    if (!skip) {
      const double v = sqrt(s);
      scal[SC_NORM] = v;
      mb[SC_NORM] = v;
      const int cnt = gate[1] + 1;
      gate[1] = cnt;
      const double thr = rel ? tol * scal[SC_NORM0] : tol;
      int code = 0;
      if (isnan(v) || isinf(v))
        code = 2;
      else if (!(v > thr))
        code = 1;
      else if (cnt >= cap)
        code = 3;
      if (code) gate[0] = 1;
      gate[2] = code;
      mb[MB_COUNT] = (double)cnt;
      mb[MB_CODE] = (double)code;
    }
    __threadfence_system();
    mb[MB_SEQ] = (double)seq;
  }
}

// Start of an FP64 solve decided on the device.  Pass 1: the residual-norm
// partials of the plain start (interior values from ui, fault and remote
// columns from u), of a candidate start (when cand is given) and of the
// boundary data, each over the same decomposition and in the same order as
// k_resnorm_partial and k_bdres_partial (8 blocks), so every norm has the
// same bits as when those kernels compute it.  partial[0..nb):
// plain start; partial[NORM_BLOCKS..+8): boundary data;
// partial[2*NORM_BLOCKS..+nb): candidate (nb = gridDim.x).
constexpr int BDRES_BLOCKS = 8;

template <bool CAND>
__global__ __launch_bounds__(NORM_THREADS) void k_start_partials(
    const double *__restrict__ ui, const double *__restrict__ u,
    const double *__restrict__ cand, const double *__restrict__ cx,
    const double *__restrict__ cz, int N, int K, int P, double DH2, int bd,
    double *__restrict__ partial, const int *__restrict__ spec, int go) {
  // a start launched ahead: only if what it waits on went through (spec[go]:
  // the solve before it converged, or its attempt was accepted)
  if (spec && spec[go] != 1) return;
  __shared__ double sh[3][NORM_THREADS];
  double acc = 0.0, accc = 0.0, accb = 0.0;
  // each thread's points of a row are evaluated together (their loads in
  // flight at once), then added in order
  constexpr int UNR = 4;
  for (int j = blockIdx.x; j < N; j += gridDim.x) {
    const int jS = (j == 0) ? 1 : j - 1;
    const int jN = (j == N - 1) ? N - 2 : j + 1;
    for (int k0 = 1 + threadIdx.x; k0 <= K - 2; k0 += UNR * NORM_THREADS) {
      double r[UNR], rc[UNR];
      // points past the row end evaluate the last interior point instead
      // (not added), so every load is issued before the first use
#pragma unroll
      for (int q = 0; q < UNR; q++) {
        const int k = min(k0 + q * NORM_THREADS, K - 2);
        r[q] = residual_at2(ui, u, cx, cz, j, jS, jN, k, K, P, DH2);
        rc[q] = CAND ? residual_at(cand, cx, cz, j, jS, jN, k, P, DH2) : 0.0;
      }
#pragma unroll
      for (int q = 0; q < UNR; q++) {
        if (k0 + q * NORM_THREADS <= K - 2) {
          acc += r[q] * r[q];
          if (CAND) accc += rc[q] * rc[q];
        }
      }
    }
  }
  if (bd && blockIdx.x < BDRES_BLOCKS) {
    for (int j = blockIdx.x * NORM_THREADS + threadIdx.x; j < N;
         j += BDRES_BLOCKS * NORM_THREADS) {
      const size_t row = (size_t)j * P;
      for (int pass = 0; pass < 2; pass++) {
        const int k = pass == 0 ? 1 : K - 2;
        if (pass == 1 && K - 2 == 1) break;
        const double uW = (k - 1 == 0) ? u[row] : 0.0;
        const double uE = (k + 1 == K - 1) ? u[row + K - 1] : 0.0;
        const double m1 = cx[row + k], m2 = cx[row + k + 1];
        const double r = 0.0 - (m1 * uW + m2 * uE) / DH2;
        accb += r * r;
      }
    }
  }
  // the three block sums (each the order of block_sum_256)
  double v3[3] = {acc, accc, accb};
  block_sums<3>(v3, &sh[0][0]);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = v3[0];
    if (CAND) partial[2 * NORM_BLOCKS + blockIdx.x] = v3[1];
    if (bd && blockIdx.x < BDRES_BLOCKS) partial[NORM_BLOCKS + blockIdx.x] = v3[2];
  }
}

// Pass 2 (one block): the norms (sums as k_norm_final), the start (the
// candidate when it is finite and below the plain start's norm, the rule of
// the host path), the threshold and the gate of the first batch: gate[0]
// stop, gate[1] = 0 V-cycles, gate[2] outcome (1 when the start already
// meets the threshold, 3 when no V-cycle is allowed), gate[4] candidate
// used, gate[5] candidate given but not used.  Counts as one step of the
// mailbox sequence.  With spec (a start launched ahead, behind a solve still
// in flight) and that solve not converged (gate[2] != 1), the start is
// abandoned: gate[0] = 1 makes the batch behind it skip, and the mailbox
// step carries outcome 4.
__global__ __launch_bounds__(NORM_THREADS) void k_start_gate(
    const double *__restrict__ partial, int nb, int cand, int rel, double tol, int cap,
    double *__restrict__ scal, double *__restrict__ mb, int *__restrict__ gate, int spec) {
  // spec: 0, or the gate word a start launched ahead waits on (2 or 7)
  if (spec && gate[spec] != 1) {
    if (threadIdx.x == 0) {
      gate[0] = 1;
      const int seq = gate[3] + 1;
      gate[3] = seq;
      mb[MB_CODE] = 4.0;
      __threadfence_system();
      mb[MB_SEQ] = (double)seq;
    }
    return;
  }
  // the three sums of k_norm_final's order, in one tree
  __shared__ double sh[3][NORM_THREADS];
  const int tid = threadIdx.x;
  double a = 0.0, b = 0.0, cc = 0.0;
  for (int i = tid; i < nb; i += NORM_THREADS) a += partial[i];
  if (rel)
    for (int i = tid; i < BDRES_BLOCKS; i += NORM_THREADS) b += partial[NORM_BLOCKS + i];
  if (cand)
    for (int i = tid; i < nb; i += NORM_THREADS) cc += partial[2 * NORM_BLOCKS + i];
  double v3[3] = {a, b, cc};
  block_sums<3>(v3, &sh[0][0]);
  const double su = v3[0], sb = v3[1], sc = v3[2];
  if (threadIdx.x == 0) {
    const double du = sqrt(su), dc = sqrt(sc);
    const double r0 = rel ? sqrt(sb) : du;
    const bool use = cand && !(isnan(dc) || isinf(dc)) && dc < du;
    const double d = use ? dc : du;
    const double thr = rel ? tol * r0 : tol;
    int code = 0;
    if (!(d > thr))
      code = 1;
    else if (cap < 1)
      code = 3;
    scal[SC_NORM] = d;
    scal[SC_NORM0] = r0;
    gate[0] = code ? 1 : 0;
    gate[1] = 0;
    gate[2] = code;
    gate[4] = use ? 1 : 0;
    gate[5] = (cand && !use) ? 1 : 0;
    mb[SC_NORM] = d;
    mb[SC_NORM0] = r0;
    if (cand) {
      scal[SC_NORM_PRED] = dc;
      mb[SC_NORM_PRED] = dc;
    }
    mb[MB_START] = d;
    mb[MB_USED] = use ? 1.0 : 0.0;
    mb[MB_COUNT] = 0.0;
    mb[MB_CODE] = (double)code;
    const int seq = gate[3] + 1;
    gate[3] = seq;
    __threadfence_system();
    mb[MB_SEQ] = (double)seq;
  }
}

// A solve that needs more V-cycles after a start was launched ahead of its
// end: that start was abandoned and stopped the gate; its V-cycles resume.
__global__ void k_gate_resume(int *gate) { gate[0] = 0; }

// dst = src when flag[0] (the candidate start was chosen).
__global__ void k_copy_if(double *__restrict__ dst, const double *__restrict__ src,
                          size_t n, const int *__restrict__ flag) {
  if (!flag[0]) return;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// u's interior columns (1..K-2) = src's when flag[0] (the plain start
// restored over a candidate that was not chosen).
__global__ void k_restore_if(double *__restrict__ u, const double *__restrict__ src,
                             int N, int K, int P, const int *__restrict__ flag) {
  if (!flag[0]) return;
  for (int j = blockIdx.x; j < N; j += gridDim.x)
    for (int k = 1 + threadIdx.x; k <= K - 2; k += blockDim.x)
      u[(size_t)j * P + k] = src[(size_t)j * P + k];
}

// IR update u += e (interior columns), FP64 += FP32.
__global__ __launch_bounds__(BX *BY) void k_add_correction(
    double *__restrict__ u, const float *__restrict__ e, int N, int K, int P) {
  const int j = blockIdx.y * BY + threadIdx.y;
  const int k = blockIdx.x * BX + threadIdx.x + 1;
  if (j >= N || k > K - 2) return;
  const size_t c = (size_t)j * P + k;
  u[c] += (double)e[c];
}

// Operator coefficients of one level from the doubled mu_equi array.  Level
// 0 reads the equivalent-medium faces (odeSolver.cpp:785-789); coarse levels
// read nodal mu at the fine node under each coarse-face midpoint
// (odeSolver.cpp:823-827).  On coarse levels the first and last rows of cz
// repeat their neighbour, reproducing the reference's use of one vertical
// coefficient twice on the Neumann rows (830-847).
__global__ void k_extract_coef(const double *__restrict__ me, int nx, int N,
                               int K, int P, int level, int fc,
                               double *__restrict__ cx,
                               double *__restrict__ cz) {
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (j > N || k >= K) return;
  const long W2 = 2L * nx - 1;
  if (j < N && k >= 1) {
    long ix;
    if (level == 0)
      ix = ((long)(k + 1) * 2 - 1) + ((long)(j + 1) * 2) * W2;
    else
      ix = ((long)((k * 2 - 1) * fc / 2) + 1) * 2 + ((long)(j * fc + 1) * 2) * W2;
    cx[(size_t)j * P + k] = me[ix];
  }
  if (k >= 1 && k <= K - 2) {
    long iz;
    if (level == 0) {
      iz = (long)(k + 1) * 2 + ((long)(j + 1) * 2 - 1) * W2;
    } else {
      const int jj = (j == 0) ? 1 : ((j == N) ? N - 1 : j);
      iz = (long)(k * fc + 1) * 2 + ((long)((jj * 2 - 1) * fc / 2) + 1) * 2 * W2;
    }
    cz[(size_t)j * P + k] = me[iz];
  }
}

__global__ void k_to_float(const double *__restrict__ a, float *__restrict__ b,
                           size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------

static inline dim3 grid_cols(int ncols, int nrows) {
  return dim3((unsigned)((ncols + BX - 1) / BX), (unsigned)((nrows + BY - 1) / BY));
}

template <typename T, bool EXACT>
static void smooth(qdeq_ctx *c, LevelBufs<T> &B, const LevelDims &d, int sweeps,
                   bool has_f) {
  const dim3 blk(BX, BY);
  const dim3 grd = grid_cols((d.K + 1) / 2, d.N);
  const T DH2 = (T)(d.DH * d.DH);
  const T relax = (T)c->cfg.relax;
  for (int it = 0; it < sweeps; it++) {
    for (int col = 0; col < 2; col++) {
      if (has_f)
        hipLaunchKernelGGL((k_rbgs<T, true, EXACT>), grd, blk, 0, c->stream,
                           B.u, B.f, B.cx, B.cz, d.N, d.K, d.P, DH2, relax, col);
      else
        hipLaunchKernelGGL((k_rbgs<T, false, EXACT>), grd, blk, 0, c->stream,
                           B.u, B.f, B.cx, B.cz, d.N, d.K, d.P, DH2, relax, col);
    }
  }
}

static void coarse_direct(qdeq_ctx *c, LevelBufs<double> &B) {
  launch_coarse_direct_fp64(c, B.u, B.f);
}
static void coarse_direct(qdeq_ctx *c, LevelBufs<float> &B) {
  launch_coarse_direct_fp32(c, B.u, B.f);
}

// vcycle (odeSolver.cpp:1101-1147) on buffers B (FP64 or FP32 hierarchy).
template <typename T, bool EXACT>
static void vcycle(qdeq_ctx *c, LevelBufs<T> *B, int level, bool level0_has_f) {
  const LevelDims &d = c->lv[level];
  const bool ref = c->cfg.cycle == QDEQ_CYCLE_REFERENCE;
  const bool has_f = level > 0 || level0_has_f;
  if (level >= c->L - 1) {
    if (!ref && c->coarse_n > 0)
      coarse_direct(c, B[level]);
    else
      smooth<T, EXACT>(c, B[level], d, c->cfg.n_coarse, has_f);
    return;
  }
  smooth<T, EXACT>(c, B[level], d, c->cfg.n_pre, has_f);
  const dim3 blk(BX, BY);
  const T DH2 = (T)(d.DH * d.DH);
  if (has_f)
    hipLaunchKernelGGL((k_residual<T, true>), grid_cols(d.K - 2, d.N), blk, 0,
                       c->stream, B[level].r, B[level].u, B[level].f,
                       B[level].cx, B[level].cz, d.N, d.K, d.P, DH2);
  else
    hipLaunchKernelGGL((k_residual<T, false>), grid_cols(d.K - 2, d.N), blk, 0,
                       c->stream, B[level].r, B[level].u, B[level].f,
                       B[level].cx, B[level].cz, d.N, d.K, d.P, DH2);
  const LevelDims &dc = c->lv[level + 1];
  hipLaunchKernelGGL((k_restrict<T>), grid_cols(dc.K - 2, dc.N), blk, 0,
                     c->stream, B[level + 1].f, B[level + 1].u, B[level].r,
                     dc.N, dc.K, dc.P, d.P);
  vcycle<T, EXACT>(c, B, level + 1, level0_has_f);
  if (ref)
    hipLaunchKernelGGL((k_prolong<T, false>), grid_cols(dc.K - 1, dc.N - 1),
                       blk, 0, c->stream, B[level].u, B[level + 1].u, dc.N,
                       dc.K, dc.P, d.P);
  else
    hipLaunchKernelGGL((k_prolong<T, true>), grid_cols(dc.K - 1, dc.N), blk, 0,
                       c->stream, B[level].u, B[level + 1].u, dc.N, dc.K, dc.P,
                       d.P);
  smooth<T, EXACT>(c, B[level], d, c->cfg.n_post, has_f);
}

// Stop test of a V-cycle launched inside a batch (gate_ptr set), or the
// plain norm of the initial check.
static void launch_norm_final(qdeq_ctx *c, int nparts) {
  if (c->gate_ptr && c->skip_gate) return;  // the next V-cycle runs this test
  if (c->gate_ptr)
    hipLaunchKernelGGL(k_norm_gate, dim3(1), dim3(NORM_THREADS), 0, c->stream,
                       c->partial, nparts, c->scal, mb_dev(c), c->gate_ptr,
                       c->cfg.tol_mode == QDEQ_TOL_REL ? 1 : 0, c->cfg.tol,
                       c->cfg.max_vcycles);
  else
    hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(NORM_THREADS), 0, c->stream,
                       c->partial, nparts, c->scal + SC_NORM, mb_dev(c) + SC_NORM);
}

// Level-0 residual norm into scal[SC_NORM] (and r32/e32 for IR).
static void launch_resnorm(qdeq_ctx *c, bool store32) {
  const LevelDims &d = c->lv[0];
  const double DH2 = d.DH * d.DH;
  if (store32)
    hipLaunchKernelGGL((k_resnorm_partial<true>), dim3(norm_blocks(c)),
                       dim3(NORM_THREADS), 0, c->stream, c->d[0].u, c->d[0].cx,
                       c->d[0].cz, d.N, d.K, d.P, DH2, c->partial, c->s[0].f,
                       c->s[0].u, (const int *)c->gate_ptr);
  else
    hipLaunchKernelGGL((k_resnorm_partial<false>), dim3(norm_blocks(c)),
                       dim3(NORM_THREADS), 0, c->stream, c->d[0].u, c->d[0].cx,
                       c->d[0].cz, d.N, d.K, d.P, DH2, c->partial,
                       (float *)nullptr, (float *)nullptr, (const int *)c->gate_ptr);
  launch_norm_final(c, norm_blocks(c));
}

// Residual norm of an arbitrary level-0 displacement u (same operator and
// summation as the stop test) into scal[slot]; partials at partial + off.
void launch_resnorm_of(qdeq_ctx *c, const double *u, int off, int slot) {
  const LevelDims &d = c->lv[0];
  hipLaunchKernelGGL((k_resnorm_partial<false>), dim3(norm_blocks(c)),
                     dim3(NORM_THREADS), 0, c->stream, u, c->d[0].cx, c->d[0].cz,
                     d.N, d.K, d.P, d.DH * d.DH, c->partial + off,
                     (float *)nullptr, (float *)nullptr, (const int *)nullptr);
  hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(NORM_THREADS), 0, c->stream,
                     c->partial + off, norm_blocks(c), c->scal + slot,
                     mb_dev(c) + slot);
}

static void launch_bdres(qdeq_ctx *c) {
  const LevelDims &d = c->lv[0];
  const int nb = 8;
  hipLaunchKernelGGL(k_bdres_partial, dim3(nb), dim3(NORM_THREADS), 0,
                     c->stream, c->d[0].u, c->d[0].cx, d.N, d.K, d.P,
                     d.DH * d.DH, c->partial + NORM_BLOCKS);
  hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(NORM_THREADS), 0, c->stream,
                     c->partial + NORM_BLOCKS, nb, c->scal + SC_NORM0,
                     mb_dev(c) + SC_NORM0);
}

// Whether solves of this context decide their start on the device: tiled
// solves (FP64, and FP32-IR without a candidate start), unless
// QDEQ_HOST_START=1.  The choices and norms are the same either way; the
// host then waits only for the V-cycles.
bool device_start(const qdeq_ctx *c) { return !c->host_start && use_tiled(c); }

// The same for a solve with a candidate start (projection or learned
// guess): FP64 only; FP32-IR candidates always take the host path, which
// expects the candidate in r and its norm in the mailbox.
bool device_start_cand(const qdeq_ctx *c) {
  return device_start(c) && c->cfg.precision != QDEQ_PREC_FP32_IR;
}

// Whether this context runs the fast cycle with the tiled kernels.
bool use_tiled(const qdeq_ctx *c) {
  return c->cfg.cycle == QDEQ_CYCLE_FAST && c->tiled &&
         tail_fits(c, c->cfg.precision != QDEQ_PREC_FP32_IR);
}

// The launch sequence of one solver iteration, followed by the residual norm.
// FP64 cycles always use the reference's update formula (bit-comparable to
// the oracle); the FP32 correction cycle uses a plain division.
static void launch_iteration(qdeq_ctx *c) {
  const bool tiled = use_tiled(c);
  if (c->cfg.precision == QDEQ_PREC_FP32_IR) {
    if (tiled) {
      vcycle_tiled_fp32_ir(c);
    } else {
      vcycle<float, false>(c, c->s, 0, true);
      const LevelDims &d = c->lv[0];
      hipLaunchKernelGGL(k_add_correction, grid_cols(d.K - 2, d.N), dim3(BX, BY),
                         0, c->stream, c->d[0].u, c->s[0].u, d.N, d.K, d.P);
    }
    launch_resnorm(c, true);
  } else if (tiled) {
    int nparts = 0;
    vcycle_tiled_fp64(c, c->partial, &nparts);
    launch_norm_final(c, nparts);
  } else {
    vcycle<double, true>(c, c->d, 0, false);
    launch_resnorm(c, false);
  }
}

void mg_destroy_graphs(qdeq_ctx *c) {
  if (c->g_cycle) (void)hipGraphExecDestroy(c->g_cycle);
  c->g_cycle = nullptr;
  c->g_cycle_mode = -1;
  for (auto &kv : c->g_stage) (void)hipGraphExecDestroy(kv.second);
  c->g_stage.clear();
}

static int ensure_graph(qdeq_ctx *c) {
  const int mode = c->cfg.precision * 2 + c->cfg.cycle + (c->tiled ? 4 : 0) +
                   64 * (c->tile_tx * 1024 + c->tile_ty);
  if (c->g_cycle && c->g_cycle_mode == mode) return QDEQ_OK;
  mg_destroy_graphs(c);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    c->gate_ptr = c->gate;
    launch_iteration(c);
    c->gate_ptr = nullptr;
    e = hipStreamEndCapture(c->stream, &g);
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&c->g_cycle, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  if (e != hipSuccess)
    return set_error(c, QDEQ_ERR_HIP,
                     std::string("V-cycle graph capture: ") + hipGetErrorString(e));
  c->g_cycle_mode = mode;
  return QDEQ_OK;
}

// Mailbox: k_norm_final writes each norm also into host-mapped memory, so
// the host reads it by polling instead of a copy and a stream sync.  A slot
// is armed with a value no norm takes (norms are >= 0 or NaN) before the
// launches that write it.
constexpr double kMboxEmpty = -1.0;
constexpr long long kTimingEvery = 8;

void mbox_arm(qdeq_ctx *c, int slot) {
  reinterpret_cast<volatile double *>(mb_host(c))[slot] = kMboxEmpty;
}

int mbox_wait(qdeq_ctx *c, int slot) {
  volatile double *p = reinterpret_cast<volatile double *>(mb_host(c)) + slot;
  long spins = 0;
  std::chrono::steady_clock::time_point idle_since{};
  bool idle = false;
  while (*p == kMboxEmpty) {
    if ((++spins & 1023) != 0) continue;
    // surface launch failures instead of spinning forever
    const hipError_t e = hipStreamQuery(c->stream);
    if (e != hipSuccess && e != hipErrorNotReady)
      return set_error(c, QDEQ_ERR_HIP,
                       std::string("waiting for a norm: ") + hipGetErrorString(e));
    if (e == hipSuccess) {
      // the stream is idle: the value lands shortly or was never written
      const auto now = std::chrono::steady_clock::now();
      if (!idle) {
        idle = true;
        idle_since = now;
      } else if (now - idle_since > std::chrono::milliseconds(200)) {
        return set_error(c, QDEQ_ERR_STATE, "norm mailbox was not written");
      }
    }
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  return QDEQ_OK;
}

// multigrid (odeSolver.cpp:652-720) on the device-resident level-0 u whose
// Dirichlet columns are already set.  Iterates while ||r||_2 > threshold.
// Non-blocking check of a mailbox slot: *ready = 1 once written.  Launch
// failures surface through an occasional hipStreamQuery, and a slot still
// empty 200 ms after the stream went idle is an error.
static int mbox_check(qdeq_ctx *c, int slot, bool block, int *ready) {
  if (block) {
    const int st = mbox_wait(c, slot);
    *ready = st == QDEQ_OK;
    return st;
  }
  volatile double *p = reinterpret_cast<volatile double *>(mb_host(c)) + slot;
  if (*p != kMboxEmpty) {
    std::atomic_thread_fence(std::memory_order_acquire);
    c->mb_idle = false;
    *ready = 1;
    return QDEQ_OK;
  }
  *ready = 0;
  if ((++c->mb_polls & 255) != 0) return QDEQ_OK;
  const hipError_t e = hipStreamQuery(c->stream);
  if (e != hipSuccess && e != hipErrorNotReady)
    return set_error(c, QDEQ_ERR_HIP,
                     std::string("waiting for a norm: ") + hipGetErrorString(e));
  if (e == hipSuccess) {
    const auto now = std::chrono::steady_clock::now();
    if (!c->mb_idle) {
      c->mb_idle = true;
      c->mb_idle_since = now;
    } else if (now - c->mb_idle_since > std::chrono::milliseconds(200)) {
      c->mb_idle = false;
      return set_error(c, QDEQ_ERR_STATE, "norm mailbox was not written");
    }
  }
  return QDEQ_OK;
}

// multigrid (odeSolver.cpp:652-720) on the device-resident level-0 u whose
// Dirichlet columns are already set, as a resumable machine: mg_begin
// enqueues the initial residual norms; mg_poll consumes each norm as it
// arrives and enqueues the next V-cycle while ||r||_2 > threshold, without
// blocking unless asked to.  A solve runs the same launches in the same order
// whether it is polled or waited for.
static int launch_batch(qdeq_ctx *c, int k);
static int launch_batch_of(qdeq_ctx *c, MgSolve &m, int k, bool allow_sample);

// The device start of a solve without a candidate: FP64 folds it into the
// first V-cycle's down kernels when that fits (fuse), else runs the start
// partials and gate; FP32-IR runs the host path's residual and norm kernels
// (the FP64 residual kept as the FP32 right-hand side) and the gate.  One
// step of the mailbox sequence.
static void launch_plain_dev_start(qdeq_ctx *c, bool fuse) {
  const LevelDims &d = c->lv[0];
  const int rel = c->cfg.tol_mode == QDEQ_TOL_REL ? 1 : 0;
  if (c->cfg.precision == QDEQ_PREC_FP32_IR) {
    if (rel)
      hipLaunchKernelGGL(k_bdres_partial, dim3(BDRES_BLOCKS), dim3(NORM_THREADS), 0,
                         c->stream, c->d[0].u, c->d[0].cx, d.N, d.K, d.P, d.DH * d.DH,
                         c->partial + NORM_BLOCKS);
    hipLaunchKernelGGL((k_resnorm_partial<true>), dim3(norm_blocks(c)),
                       dim3(NORM_THREADS), 0, c->stream, c->d[0].u, c->d[0].cx,
                       c->d[0].cz, d.N, d.K, d.P, d.DH * d.DH, c->partial, c->s[0].f,
                       c->s[0].u, (const int *)nullptr);
  } else if (fuse) {
    c->fuse_start_next = true;
    return;
  } else {
    hipLaunchKernelGGL(k_start_partials<false>, dim3(norm_blocks(c)), dim3(NORM_THREADS),
                       0, c->stream, c->d[0].u, c->d[0].u, (const double *)nullptr,
                       c->d[0].cx, c->d[0].cz, d.N, d.K, d.P, d.DH * d.DH, rel,
                       c->partial, (const int *)(c->spec_launch ? c->gate : nullptr),
                       c->spec_go);
  }
  hipLaunchKernelGGL(k_start_gate, dim3(1), dim3(NORM_THREADS), 0, c->stream,
                     c->partial, norm_blocks(c), 0, rel, c->cfg.tol, c->cfg.max_vcycles,
                     c->scal, mb_dev(c), c->gate, c->spec_launch ? c->spec_go : 0);
}

// V-cycles of the first batch of a solve: what the same RK43 stage needed in
// the previous step (c->last_it_key), else what the last solve needed.
static int first_batch(const qdeq_ctx *c) {
  const int key = c->solve_key;
  const int pred = (key >= 0 && c->last_it_key[key] >= 0) ? c->last_it_key[key]
                                                           : c->last_it;
  return std::min(std::max(pred, 1), 6);
}

int mg_begin(qdeq_ctx *c) {
  int st = ensure_graph(c);
  if (st != QDEQ_OK) return st;
  MgSolve &m = c->mg;
  // no other solve is in flight: mailbox page 0
  m.page = 0;
  c->mb_page = 0;
  m.ir = c->cfg.precision == QDEQ_PREC_FP32_IR;
  // a predicted start (learned guess) was enqueued before this solve: its
  // candidate sits in the level-0 r buffer, its residual norm in
  // scal[SC_NORM_PRED] (and the mailbox slot, armed before the prediction)
  m.cand = c->cand_pending;
  c->cand_pending = false;
  m.it = 0;
  m.sample = false;
  m.launched = 0;
  m.dev_start = m.cand ? device_start_cand(c) : device_start(c);
  m.start_read = false;
  if (m.dev_start && m.ir) {
    // FP32-IR: the start decided on the device by the host path's kernels
    // and rule (same sums, same test)
    launch_plain_dev_start(c, false);
    c->seq_issued += 1;
    QDEQ_HIP(c, hipGetLastError());
    int st = launch_batch(c, first_batch(c));
    if (st != QDEQ_OK) return st;
    m.phase = 2;
    return QDEQ_OK;
  }
  if (m.dev_start) {
    // norms, start choice and stop test on the device; the first batch is
    // launched at once, sized from the last solve
    // the candidate is either written over u's interior with the plain
    // start's interior kept at plain_interior (the projection start), or in
    // r (the learned guess)
    const LevelDims &d = c->lv[0];
    const int rel = c->cfg.tol_mode == QDEQ_TOL_REL ? 1 : 0;
    const bool in_u = m.cand && c->cand_in_u;
    c->cand_in_u = false;
    const double *plain = in_u ? c->plain_interior : c->d[0].u;
    const double *cand = !m.cand ? nullptr : (in_u ? c->d[0].u : c->d[0].r);
    // a plain start launched kernel by kernel: its norms and test ride on
    // the first V-cycle's down kernels (the same decisions, no start launches)
    if (!cand) {
      launch_plain_dev_start(c, launch_direct(c) && fused_start_fits(c));
    } else {
      hipLaunchKernelGGL(k_start_partials<true>, dim3(norm_blocks(c)), dim3(NORM_THREADS), 0,
                         c->stream, plain, c->d[0].u, cand, c->d[0].cx, c->d[0].cz,
                         d.N, d.K, d.P, d.DH * d.DH, rel, c->partial, (const int *)nullptr,
                         2);
      hipLaunchKernelGGL(k_start_gate, dim3(1), dim3(NORM_THREADS), 0, c->stream,
                         c->partial, norm_blocks(c), 1, rel, c->cfg.tol,
                         c->cfg.max_vcycles, c->scal, mb_dev(c), c->gate, 0);
    }
    if (in_u) {
      // the plain start back over a candidate that was not chosen
      hipLaunchKernelGGL(k_restore_if, dim3(std::min(d.N, 1024)), dim3(256), 0,
                         c->stream, c->d[0].u, plain, d.N, d.K, d.P, c->gate + 5);
    } else if (m.cand) {
      // the chosen candidate from r into u
      const size_t n = (size_t)d.N * d.P;
      const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 1024);
      hipLaunchKernelGGL(k_copy_if, dim3(blocks), dim3(256), 0, c->stream,
                         c->d[0].u, c->d[0].r, n, c->gate + 4);
    }
    c->seq_issued += 1;
    QDEQ_HIP(c, hipGetLastError());
    int st = launch_batch(c, first_batch(c));
    if (st != QDEQ_OK) return st;
    m.phase = 2;
    return QDEQ_OK;
  }
  mbox_arm(c, SC_NORM);
  mbox_arm(c, SC_NORM0);
  if (c->cfg.tol_mode == QDEQ_TOL_REL) launch_bdres(c);
  launch_resnorm(c, m.ir);
  QDEQ_HIP(c, hipGetLastError());
  m.phase = 1;
  return QDEQ_OK;
}

// Whole RK43 stages as graph replays (where V-cycles are replays anyway):
// the stage's state kernel (kind 0: k_stage of stage s; kind 1: the fault
// update of stage s - 1 with the state of stage s), the plain device start
// and the first batch of V-cycles are one captured graph per (kind, stage,
// batch size), replayed for every solve of that stage; the same kernels with
// the same arguments as the direct launches (t and dt are read on the
// device), so the results do not change.  Later batches of the solve replay
// the one-V-cycle graph.
bool stage_graphs(const qdeq_ctx *c) {
  return !c->stage_graphs_off && !launch_direct(c) && device_start(c) && c->timing == 0 &&
         c->proj_slots == 0 &&
         (c->guess == nullptr || c->guess_mode == QDEQ_GUESS_OFF);
}

int mg_begin_stage(qdeq_ctx *c, int kind, int s) {
  int st = ensure_graph(c);
  if (st != QDEQ_OK) return st;
  MgSolve &m = c->mg;
  m.page = 0;
  c->mb_page = 0;
  m.ir = c->cfg.precision == QDEQ_PREC_FP32_IR;
  m.cand = false;
  c->cand_pending = false;
  m.it = 0;
  m.sample = false;
  m.dev_start = true;
  m.start_read = false;
  const int B = first_batch(c);
  const int gk = (kind * 8 + s) * 8 + B;
  auto it = c->g_stage.find(gk);
  hipGraphExec_t ex = nullptr;
  if (it != c->g_stage.end()) {
    ex = it->second;
  } else {
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      if (kind == 1)
        launch_update_stage(c, s - 1);
      else
        launch_stage(c, s);
      launch_plain_dev_start(c, !m.ir && fused_start_fits(c));
      c->gate_ptr = c->gate;
      for (int i = 0; i < B; i++) launch_iteration(c);
      c->gate_ptr = nullptr;
      c->fuse_start_next = false;
      e = hipStreamEndCapture(c->stream, &g);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (e != hipSuccess)
      return set_error(c, QDEQ_ERR_HIP,
                       std::string("stage graph capture: ") + hipGetErrorString(e));
    c->g_stage[gk] = ex;
  }
  QDEQ_HIP(c, hipGraphLaunch(ex, c->stream));
  c->seq_issued += 1 + B;
  m.launched = B;
  m.seq_target = c->seq_issued;
  m.phase = 2;
  return QDEQ_OK;
}

// The next FP64 solve launched ahead, behind the current solve's launched
// V-cycles (its fault update already queued by the caller): the plain
// device start and a first batch, into c->mg_next, on the other mailbox
// page.  Its kernels go on only if the current solve ends converged with
// the work already queued; if the current solve needs more V-cycles the
// start abandons itself (outcome 4) and mg_poll resumes the gate before the
// continuation.  The same kernels and arguments as mg_begin's, so a solve
// launched ahead and taken over computes what it would have computed.
int mg_begin_spec(qdeq_ctx *c, int go) {
  MgSolve &m = c->mg_next;
  m = MgSolve{};
  m.dev_start = true;
  m.page = 1 - c->mg.page;
  c->mb_page = m.page;
  c->spec_launch = true;
  c->spec_go = go;
  launch_plain_dev_start(c, fused_start_fits(c));
  c->seq_issued += 1;
  const int st = launch_batch_of(c, m, first_batch(c), true);
  c->spec_launch = false;
  c->spec_go = 2;
  c->fuse_start_next = false;
  c->mb_page = c->mg.page;
  if (st != QDEQ_OK) return st;
  QDEQ_HIP(c, hipGetLastError());
  m.phase = 2;
  c->spec_pending = true;
  return QDEQ_OK;
}

static int mg_finish(qdeq_ctx *c, int st) {
  MgSolve &m = c->mg;
  c->vcycles += m.it;
  c->last_it = m.it;
  if (c->solve_key >= 0) c->last_it_key[c->solve_key] = m.it;
  c->solve_key = -1;
  if (m.ir) c->ir_sweeps += m.it;
  c->solves += 1;
  // contraction per V-cycle seen by this solve (for the next batch sizes)
  if (st == QDEQ_OK && m.it > 0 && m.delta0 > 0 && m.delta > 0) {
    const double r = std::pow(m.delta / m.delta0, 1.0 / m.it);
    if (r > 0 && r < 1) c->rho_est = 0.7 * c->rho_est + 0.3 * r;
  }
  m.phase = 0;
  return st;
}

// Non-blocking (or blocking) check that the batch's last V-cycle has been
// seen by the device: mb[MB_SEQ] reaches the batch's sequence number.
static int seq_check(qdeq_ctx *c, long long target, bool block, int *ready) {
  volatile double *p = reinterpret_cast<volatile double *>(mb_host(c)) + MB_SEQ;
  for (;;) {
    if ((long long)*p >= target) {
      std::atomic_thread_fence(std::memory_order_acquire);
      c->mb_idle = false;
      *ready = 1;
      return QDEQ_OK;
    }
    *ready = 0;
    if ((++c->mb_polls & 255) == 0) {
      const hipError_t e = hipStreamQuery(c->stream);
      if (e != hipSuccess && e != hipErrorNotReady)
        return set_error(c, QDEQ_ERR_HIP,
                         std::string("waiting for V-cycles: ") + hipGetErrorString(e));
      if (e == hipSuccess) {
        const auto now = std::chrono::steady_clock::now();
        if (!c->mb_idle) {
          c->mb_idle = true;
          c->mb_idle_since = now;
        } else if (now - c->mb_idle_since > std::chrono::milliseconds(200)) {
          c->mb_idle = false;
          return set_error(c, QDEQ_ERR_STATE, "V-cycle sequence was not written");
        }
      }
    }
    if (!block) return QDEQ_OK;
  }
}

// Launches up to k V-cycles back to back; each skips once the solve has
// converged (or failed), so launching more than needed changes nothing.
static int launch_batch(qdeq_ctx *c, int k) { return launch_batch_of(c, c->mg, k, true); }

// The batch of solve m (the current one, or one launched ahead; each
// mailbox page has its own pair of timing events, so either may hold a
// sample).
static int launch_batch_of(qdeq_ctx *c, MgSolve &m, int k, bool allow_sample) {
  m.sample = false;
  // the stop test of every V-cycle but the batch's last (and one that could
  // reach the V-cycle cap) runs in the next V-cycle's level-0 down kernel
  const bool fold = fold_fits(c);
  bool folded = false;
  struct FoldReset {  // also on an error return: no later launch folds
    qdeq_ctx *c;
    ~FoldReset() { c->fold_next = c->skip_gate = false; }
  } fold_reset{c};
  for (int i = 0; i < k; i++) {
    // kernel timing samples every kTimingEvery-th V-cycle: that one is
    // launched directly (same kernels and arguments as the graph) with the
    // events recorded around the timed kernel
    const bool sample = allow_sample && c->timing != 0 && c->timing != QDEQ_TIME_FAULT &&
                        use_tiled(c) && (c->timed_seq++ % kTimingEvery) == 0;
    m.launched++;
    c->fold_next = folded;
    c->skip_gate = fold && i + 1 < k && m.launched < c->cfg.max_vcycles;
    folded = c->skip_gate;
    if (folded) c->folded_tests++;
    if (sample) {
      c->emit_events = true;
      c->gate_ptr = c->gate;
      launch_iteration(c);
      c->gate_ptr = nullptr;
      c->emit_events = false;
      QDEQ_HIP(c, hipGetLastError());
      m.sample = true;
      m.sample_ordinal = m.launched;
    } else if (launch_direct(c)) {
      c->gate_ptr = c->gate;
      launch_iteration(c);
      c->gate_ptr = nullptr;
      QDEQ_HIP(c, hipGetLastError());
    } else {
      QDEQ_HIP(c, hipGraphLaunch(c->g_cycle, c->stream));
    }
    c->fold_next = c->skip_gate = false;
  }
  c->seq_issued += k;
  m.seq_target = c->seq_issued;
  return QDEQ_OK;
}

// V-cycles to launch for the residual ratio left, from the observed
// contraction (the tiled cycles only; one at a time otherwise or when
// kernels are being timed).
static int batch_size(const qdeq_ctx *c, double delta, double thr) {
  if (!use_tiled(c) || (c->timing != 0 && c->timing != QDEQ_TIME_FAULT)) return 1;
  const double rho = std::min(std::max(c->rho_est, 1e-3), 0.9);
  if (!(delta > 0) || !(thr > 0)) return 1;
  const double need = std::log(thr / delta) / std::log(rho);
  int k = (int)std::ceil(need - 1e-9);
  return std::min(std::max(k, 1), 6);
}

int mg_poll(qdeq_ctx *c, bool block, int *done) {
  MgSolve &m = c->mg;
  c->mb_page = m.page;
  const double *mb = mb_host(c);
  *done = 0;
  int st = QDEQ_OK, ready;
  if (m.phase == 1) {
    if ((st = mbox_check(c, SC_NORM, block, &ready)) || !ready) goto out;
    if (c->cfg.tol_mode == QDEQ_TOL_REL &&
        ((st = mbox_check(c, SC_NORM0, block, &ready)) || !ready))
      goto out;
    if (m.cand && ((st = mbox_check(c, SC_NORM_PRED, block, &ready)) || !ready))
      goto out;
    for (int i : {SC_NORM, SC_NORM0, SC_NORM_PRED}) c->scal_host[i] = mb[i];
    m.delta = c->scal_host[SC_NORM];
    if (m.cand) {
      // used only if its residual is finite and below the previous solution's
      const double pn = c->scal_host[SC_NORM_PRED];
      if (std::isfinite(pn) && pn < m.delta) {
        const LevelDims &d = c->lv[0];
        QDEQ_HIP(c, hipMemcpyAsync(c->d[0].u, c->d[0].r,
                                   (size_t)d.N * d.P * sizeof(double),
                                   hipMemcpyDeviceToDevice, c->stream));
        if (m.ir) launch_resnorm(c, true);  // FP32 residual of the new start
        m.delta = pn;
        c->g_accepted++;
      }
    }
    m.delta0 = m.delta;
    m.r0 = c->cfg.tol_mode == QDEQ_TOL_REL ? c->scal_host[SC_NORM0] : m.delta;
    m.thr = c->cfg.tol_mode == QDEQ_TOL_REL ? c->cfg.tol * m.r0 : c->cfg.tol;
    if (!(m.delta > m.thr)) {
      *done = 1;
      return mg_finish(c, QDEQ_OK);
    }
    if (c->cfg.max_vcycles < 1) {
      *done = 1;
      return mg_finish(c, set_error(c, QDEQ_ERR_MG_NOT_CONVERGED,
                                    "multigrid: V-cycle cap reached"));
    }
    QDEQ_HIP(c, hipMemsetAsync(c->gate, 0, 3 * sizeof(int), c->stream));
    if ((st = launch_batch(c, batch_size(c, m.delta, m.thr)))) goto out;
    m.phase = 2;
    if (!block) return QDEQ_OK;
  }
  for (;;) {
    if (m.phase != 2) return set_error(c, QDEQ_ERR_STATE, "multigrid: no solve in progress");
    if ((st = seq_check(c, m.seq_target, block, &ready)) || !ready) goto out;
    if (m.sample) {
      // counted only if the timed V-cycle ran (the solve had not stopped
      // before it)
      float ms = 0.0f;
      QDEQ_HIP(c, hipEventSynchronize(c->ev_t[m.page][1]));
      if ((int)mb[MB_COUNT] >= m.sample_ordinal &&
          hipEventElapsedTime(&ms, c->ev_t[m.page][0], c->ev_t[m.page][1]) == hipSuccess) {
        c->timed_ms += ms;
        c->timed_calls += 1;
      }
      m.sample = false;
    }
    if (m.dev_start && !m.start_read) {
      // the start's norms, written before the first sequence step
      const double r0 = mb[SC_NORM0];
      m.delta0 = mb[MB_START];
      m.r0 = c->cfg.tol_mode == QDEQ_TOL_REL ? r0 : m.delta0;
      m.thr = c->cfg.tol_mode == QDEQ_TOL_REL ? c->cfg.tol * r0 : c->cfg.tol;
      c->scal_host[SC_NORM0] = r0;
      if (m.cand) {
        c->scal_host[SC_NORM_PRED] = mb[SC_NORM_PRED];
        if (mb[MB_USED] != 0.0) c->g_accepted++;
      }
      m.start_read = true;
    }
    const int code = (int)mb[MB_CODE];
    m.it = (int)mb[MB_COUNT];
    m.delta = c->scal_host[SC_NORM] = mb[SC_NORM];
    if (code == 2) {
      *done = 1;
      return mg_finish(c, set_error(c, QDEQ_ERR_MG_NOT_CONVERGED,
                                    "multigrid: residual is not finite"));
    }
    if (code == 3) {
      *done = 1;
      return mg_finish(c, set_error(c, QDEQ_ERR_MG_NOT_CONVERGED,
                                    "multigrid: V-cycle cap reached (residual " +
                                        std::to_string(m.delta) + " > " +
                                        std::to_string(m.thr) + ")"));
    }
    if (code == 1) {
      *done = 1;
      return mg_finish(c, QDEQ_OK);
    }
    // not there yet: another batch sized from the residual left.  A solve
    // launched ahead behind this one was abandoned and stopped the gate.
    if (c->spec_pending) {
      hipLaunchKernelGGL(k_gate_resume, dim3(1), dim3(1), 0, c->stream, c->gate);
      c->spec_pending = false;
    }
    if ((st = launch_batch(c, batch_size(c, m.delta, m.thr)))) goto out;
    if (!block) return QDEQ_OK;
  }
out:
  if (st != QDEQ_OK) {
    *done = 1;
    return mg_finish(c, st);
  }
  return QDEQ_OK;
}

int mg_solve(qdeq_ctx *c, int *vcycles, double *res0, double *res) {
  int st = mg_begin(c);
  if (st != QDEQ_OK) return st;
  int done = 0;
  while (!done) {
    st = mg_poll(c, true, &done);
    if (st != QDEQ_OK) break;
  }
  if (vcycles) *vcycles = c->mg.it;
  if (res0) *res0 = c->mg.r0;
  if (res) *res = c->mg.delta;
  return st;
}

// Extracts every level's coefficients (FP64, and FP32 copies for IR).
int mg_build_coefficients(qdeq_ctx *c) {
  for (int l = 0; l < c->L; l++) {
    const LevelDims &d = c->lv[l];
    const dim3 blk(64, 4);
    const dim3 grd((d.K + 63) / 64, (d.N + 1 + 3) / 4);
    hipLaunchKernelGGL(k_extract_coef, grd, blk, 0, c->stream, c->mu_equi,
                       c->nx, d.N, d.K, d.P, l, d.fc, c->d[l].cx, c->d[l].cz);
    const size_t ncx = (size_t)d.N * d.P, ncz = (size_t)(d.N + 1) * d.P;
    hipLaunchKernelGGL(k_to_float, dim3((unsigned)((ncx + 255) / 256)), dim3(256),
                       0, c->stream, c->d[l].cx, c->s[l].cx, ncx);
    hipLaunchKernelGGL(k_to_float, dim3((unsigned)((ncz + 255) / 256)), dim3(256),
                       0, c->stream, c->d[l].cz, c->s[l].cz, ncz);
  }
  QDEQ_HIP(c, hipGetLastError());
  QDEQ_HIP(c, hipStreamSynchronize(c->stream));
  return QDEQ_OK;
}

}  // namespace qdeq
