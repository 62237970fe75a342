// nn.hip -- kernels of the learned initial guess network: a BF16 MFMA GEMM
// with FP32 accumulation (gfx950 v_mfma_f32_16x16x32_bf16), the 1D
// convolution encoder, activation/mask epilogues with transposed copies, the
// MSE loss, minibatch gathering and Adam.  No reference counterpart: the
// reference starts every elliptic solve from the previous solution
// (odeSolver.cpp:652-720); this network predicts a better start.
//
// GEMM convention: C[M x N] = A[M x K] * B[N x K]^T with A and B row-major
// in their K dimension (both operands K-contiguous), C row-major FP32.  Every
// product of the forward and backward pass is written in this form by
// keeping transposed BF16 copies of the weights, activations and gradients.

#include "nn.h"

#include <algorithm>

namespace qdeq {
namespace nn {

typedef bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// ---- GEMM ----
// Block = WM x WN wavefronts; each wavefront computes TM x TN tiles of 16x16
// with the 16x16x32 MFMA.  K is consumed 32 at a time through a double-
// buffered LDS stage (rows padded to 40 elements = 80 bytes).  Operand
// lanes: lane l feeds row/column (l & 15) and the k-group (l >> 4) of eight
// consecutive k values, the same k-group for both operands; the result
// element r of lane l is row 4*(l >> 4) + r, column (l & 15) of the tile.
// blockIdx.z selects a K chunk [z*kchunk, (z+1)*kchunk) and the FP32 output
// slice C + z*split_stride (split-K partial sums, reduced by k_split_reduce).
template <int WM, int WN, int TM, int TN>
__global__ __launch_bounds__(64 * WM * WN) void k_gemm_nt(
    const bf16 *__restrict__ A, int lda, const bf16 *__restrict__ B, int ldb,
    float *__restrict__ C, int ldc, int kchunk, long long split_stride) {
  constexpr int BM = WM * TM * 16, BN = WN * TN * 16, BK = 32;
  constexpr int NT = 64 * WM * WN, LDK = BK + 8;
  constexpr int ACH = BM * BK / 8, BCH = BN * BK / 8;
  constexpr int AIT = (ACH + NT - 1) / NT, BIT = (BCH + NT - 1) / NT;
  __shared__ __attribute__((aligned(16))) bf16 As[2][BM * LDK];
  __shared__ __attribute__((aligned(16))) bf16 Bs[2][BN * LDK];

  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int kbeg = blockIdx.z * kchunk, kend = kbeg + kchunk;
  C += (long long)blockIdx.z * split_stride;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WN, wn = wave % WN;

  f32x4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  bf16x8 ra[AIT], rb[BIT];
  auto gload = [&](int k) {
#pragma unroll
    for (int i = 0; i < AIT; i++) {
      const int ch = tid + i * NT;
      if (ACH % NT == 0 || ch < ACH) {
        const int row = ch >> 2, kk = (ch & 3) * 8;
        ra[i] = *reinterpret_cast<const bf16x8 *>(A + (long long)(m0 + row) * lda + k + kk);
      }
    }
#pragma unroll
    for (int i = 0; i < BIT; i++) {
      const int ch = tid + i * NT;
      if (BCH % NT == 0 || ch < BCH) {
        const int row = ch >> 2, kk = (ch & 3) * 8;
        rb[i] = *reinterpret_cast<const bf16x8 *>(B + (long long)(n0 + row) * ldb + k + kk);
      }
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int i = 0; i < AIT; i++) {
      const int ch = tid + i * NT;
      if (ACH % NT == 0 || ch < ACH) {
        const int row = ch >> 2, kk = (ch & 3) * 8;
        *reinterpret_cast<bf16x8 *>(&As[buf][row * LDK + kk]) = ra[i];
      }
    }
#pragma unroll
    for (int i = 0; i < BIT; i++) {
      const int ch = tid + i * NT;
      if (BCH % NT == 0 || ch < BCH) {
        const int row = ch >> 2, kk = (ch & 3) * 8;
        *reinterpret_cast<bf16x8 *>(&Bs[buf][row * LDK + kk]) = rb[i];
      }
    }
  };

  gload(kbeg);
  sstore(0);
  __syncthreads();
  int buf = 0;
  const int ar = wm * TM * 16 + (lane & 15), br = wn * TN * 16 + (lane & 15);
  const int kg = (lane >> 4) * 8;
  for (int k = kbeg; k < kend; k += BK) {
    const bool more = k + BK < kend;
    if (more) gload(k + BK);
    bf16x8 a[TM], b[TN];
#pragma unroll
    for (int i = 0; i < TM; i++)
      a[i] = *reinterpret_cast<const bf16x8 *>(&As[buf][(ar + i * 16) * LDK + kg]);
#pragma unroll
    for (int j = 0; j < TN; j++)
      b[j] = *reinterpret_cast<const bf16x8 *>(&Bs[buf][(br + j * 16) * LDK + kg]);
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
      for (int j = 0; j < TN; j++)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    if (more) sstore(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  const int rbase = m0 + wm * TM * 16 + 4 * (lane >> 4);
  const int cbase = n0 + wn * TN * 16 + (lane & 15);
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        C[(long long)(rbase + i * 16 + r) * ldc + cbase + j * 16] = acc[i][j][r];
}

__global__ void k_split_reduce(const float *__restrict__ part, int splits,
                               long long n, float *__restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = part[i];
  for (int z = 1; z < splits; z++) s += part[(long long)z * n + i];
  out[i] = s;
}

// Tile configurations: "wide" for training (128 x 128 per block), "skinny"
// for the 16-row inference products (16 x 128 per block).
#define GEMM_WIDE 2, 2, 4, 4
#define GEMM_SKINNY 1, 4, 1, 2

bool gemm_shape_ok(int cfg, int M, int N, int K, int lda, int ldb) {
  const int BM = cfg == GEMM_CFG_WIDE ? 128 : 16;
  const int BN = 128;
  return M > 0 && N > 0 && K > 0 && M % BM == 0 && N % BN == 0 && K % 32 == 0 &&
         lda % 8 == 0 && ldb % 8 == 0 && lda >= K && ldb >= K;
}

int gemm_nt(hipStream_t st, int cfg, const bf16 *A, int lda, const bf16 *B,
            int ldb, float *C, int ldc, int M, int N, int K, float *splitbuf,
            size_t splitbuf_floats, int min_blocks) {
  if (!gemm_shape_ok(cfg, M, N, K, lda, ldb) || ldc < N) return -1;
  const int BM = cfg == GEMM_CFG_WIDE ? 128 : 16, BN = 128;
  const int gm = M / BM, gn = N / BN, nk32 = K / 32;
  const long long mn = (long long)M * N;
  // smallest split of K into equal whole chunks that gives min_blocks
  // blocks (at most 64 chunks of at least 4 k-steps; FP32 partials must fit)
  int splits = 1;
  if (ldc == N && splitbuf)
    for (int s = 2; s <= 64 && gm * gn * splits < min_blocks; s++) {
      if (nk32 % s || nk32 / s < 4) continue;
      if ((size_t)s * (size_t)mn > splitbuf_floats) break;
      splits = s;
    }
  const int kchunk = (nk32 / splits) * 32;
  float *out = splits > 1 ? splitbuf : C;
  const dim3 grid(gn, gm, splits);
  if (cfg == GEMM_CFG_WIDE)
    hipLaunchKernelGGL((k_gemm_nt<GEMM_WIDE>), grid, dim3(256), 0, st, A, lda, B,
                       ldb, out, ldc, kchunk, splits > 1 ? mn : 0);
  else
    hipLaunchKernelGGL((k_gemm_nt<GEMM_SKINNY>), grid, dim3(256), 0, st, A, lda,
                       B, ldb, out, ldc, kchunk, splits > 1 ? mn : 0);
  if (splits > 1)
    hipLaunchKernelGGL(k_split_reduce, dim3((unsigned)((mn + 255) / 256)),
                       dim3(256), 0, st, splitbuf, splits, mn, C);
  return 0;
}

// ---- epilogues ----
// out = bf16(op(Z)) and outT = its transpose (either may be null); op: RELU
// max(z,0); MASK z*(mask>0); PLAIN z.  outF (FP32, row-major) receives the
// masked value unrounded when given.  32x32 tiles through LDS.
__global__ void k_epi(const float *__restrict__ Z, int M, int N, int mode,
                      const bf16 *__restrict__ mask, bf16 *__restrict__ out,
                      bf16 *__restrict__ outT, float *__restrict__ outF) {
  __shared__ float tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
  for (int r = ty; r < 32; r += 8) {
    const int i = by + r, j = bx + tx;
    float v = 0.f;
    if (i < M && j < N) {
      const long long idx = (long long)i * N + j;
      v = Z[idx];
      if (mode == EPI_RELU)
        v = v > 0.f ? v : 0.f;
      else if (mode == EPI_MASK)
        v = (float)mask[idx] > 0.f ? v : 0.f;
      if (out) out[idx] = (bf16)v;
      if (outF) outF[idx] = v;
    }
    tile[r][tx] = v;
  }
  if (!outT) return;
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int j = bx + r, i = by + tx;
    if (i < M && j < N) outT[(long long)j * M + i] = (bf16)tile[tx][r];
  }
}

void epilogue(hipStream_t st, const float *Z, int M, int N, int mode,
              const bf16 *mask, bf16 *out, bf16 *outT, float *outF) {
  hipLaunchKernelGGL(k_epi, dim3((N + 31) / 32, (M + 31) / 32), dim3(32, 8), 0,
                     st, Z, M, N, mode, mask, out, outT, outF);
}

// ---- encoder: Conv1d(1 -> C, width T, stride S), zero padding, ReLU ----
// Output position p reads nodes S*p - (T-S)/2 + t, t < T; feature p*C + c.
__global__ void k_conv_fwd(const float *__restrict__ X, int M, int nk,
                           const float *__restrict__ Wc, int C, int T, int S,
                           int P0, bf16 *__restrict__ H0, bf16 *__restrict__ H0T) {
  const int F0 = P0 * C;
  const long long n = (long long)M * F0;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n) return;
  const int b = (int)(gid % M);  // batch fastest: coalesced transposed store
  const int f = (int)(gid / M);
  const int p = f / C, c = f % C;
  const int x0 = S * p - (T - S) / 2;
  const float *x = X + (long long)b * nk;
  float acc = 0.f;
  for (int t = 0; t < T; t++) {
    const int q = x0 + t;
    if (q >= 0 && q < nk) acc = acc + Wc[c * T + t] * x[q];
  }
  const bf16 h = (bf16)(acc > 0.f ? acc : 0.f);
  H0[(long long)b * F0 + f] = h;
  if (H0T) H0T[(long long)f * M + b] = h;
}

void conv_forward(hipStream_t st, const float *X, int M, int nk, const float *Wc,
                  int C, int T, int S, int P0, bf16 *H0, bf16 *H0T) {
  const long long n = (long long)M * P0 * C;
  hipLaunchKernelGGL(k_conv_fwd, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     st, X, M, nk, Wc, C, T, S, P0, H0, H0T);
}

// dWc[c][t] = sum_b sum_p dZ0[b][p*C+c] * x[b][S*p - (T-S)/2 + t]; one block
// per (c, t), fixed-order tree reduction.
__global__ void k_conv_bwd(const float *__restrict__ X, int M, int nk,
                           const float *__restrict__ dZ0, int C, int T, int S,
                           int P0, float *__restrict__ gWc) {
  __shared__ float red[256];
  const int c = blockIdx.x / T, t = blockIdx.x % T;
  const int F0 = P0 * C;
  float acc = 0.f;
  for (int i = threadIdx.x; i < M * P0; i += 256) {
    const int b = i / P0, p = i % P0;
    const int q = S * p - (T - S) / 2 + t;
    if (q >= 0 && q < nk)
      acc = acc + dZ0[(long long)b * F0 + p * C + c] * X[(long long)b * nk + q];
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) gWc[c * T + t] = red[0];
}

void conv_backward(hipStream_t st, const float *X, int M, int nk, const float *dZ0,
                   int C, int T, int S, int P0, float *gWc) {
  hipLaunchKernelGGL(k_conv_bwd, dim3(C * T), dim3(256), 0, st, X, M, nk, dZ0,
                     C, T, S, P0, gWc);
}

// ---- loss ----
// For rows b < B: d = Y - T; Y <- 2 d / (B * nout) (the gradient, in place);
// the rows >= B and the padding columns >= nout are zeroed.  Per-block sums
// of d^2 land in partial[blockIdx.x] (256 threads, fixed-order tree).
__global__ void k_mse(float *__restrict__ Y, const float *__restrict__ Tg, int Mb,
                      int B, int nout, int ld, double *__restrict__ partial) {
  __shared__ double red[256];
  const long long n = (long long)Mb * ld;
  const float scale = 2.0f / ((float)B * (float)nout);
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (long long)gridDim.x * 256) {
    const int b = (int)(i / ld), j = (int)(i % ld);
    float g = 0.f;
    if (b < B && j < nout) {
      const float d = Y[i] - Tg[i];
      acc += (double)d * (double)d;
      g = scale * d;
    }
    Y[i] = g;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_sum_partials(const double *__restrict__ partial, int n,
                               double scale, double *__restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s = s + partial[i];
    *out = s * scale;
  }
}

void mse(hipStream_t st, float *Y, const float *Tg, int Mb, int B, int nout,
         int ld, double *partial, double *loss_out) {
  hipLaunchKernelGGL(k_mse, dim3(MSE_BLOCKS), dim3(256), 0, st, Y, Tg, Mb, B,
                     nout, ld, partial);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, st, partial,
                     MSE_BLOCKS, 1.0 / ((double)B * (double)nout), loss_out);
}

// ---- minibatch gather: rows b < B copy pair idx[b]; rows >= B are zero ----
__global__ void k_gather(const float *__restrict__ xs, const float *__restrict__ ts,
                         const int *__restrict__ idx, int B, int Mb, int nk,
                         int ld, float *__restrict__ X, float *__restrict__ Tg) {
  const int b = blockIdx.y;
  const long long src = b < B ? idx[b] : -1;
  for (int j = blockIdx.x * 256 + threadIdx.x; j < ld; j += gridDim.x * 256) {
    Tg[(long long)b * ld + j] = src >= 0 ? ts[src * ld + j] : 0.f;
    if (j < nk) X[(long long)b * nk + j] = src >= 0 ? xs[src * nk + j] : 0.f;
  }
  (void)Mb;
}

void gather(hipStream_t st, const float *xs, const float *ts, const int *idx,
            int B, int Mb, int nk, int ld, float *X, float *Tg) {
  hipLaunchKernelGGL(k_gather, dim3(32, Mb), dim3(256), 0, st, xs, ts, idx, B, Mb,
                     nk, ld, X, Tg);
}

// ---- Adam (bias-corrected), FP32 master weights ----
__global__ void k_adam(float *__restrict__ w, float *__restrict__ m,
                       float *__restrict__ v, const float *__restrict__ g,
                       long long n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

void adam(hipStream_t st, float *w, float *m, float *v, const float *g,
          long long n, float lr, float b1, float b2, float eps, float c1,
          float c2) {
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     w, m, v, g, n, lr, b1, b2, eps, c1, c2);
}

}  // namespace nn
}  // namespace qdeq
