// nn.h -- host launchers of the learned-guess network kernels (nn.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace qdeq {
namespace nn {

typedef __bf16 bf16;

enum { GEMM_CFG_WIDE = 0, GEMM_CFG_SKINNY = 1 };
enum { EPI_PLAIN = 0, EPI_RELU = 1, EPI_MASK = 2 };
enum { MSE_BLOCKS = 512 };

// C[M x N] (FP32, row pitch ldc) = A[M x K] * B[N x K]^T, BF16 operands with
// row pitches lda, ldb (multiples of 8), FP32 accumulation.  WIDE tiles need
// M, N multiples of 128, SKINNY ones M % 16 == 0 and N % 128 == 0; K % 32 ==
// 0.  K is split over blocks (FP32 partials in splitbuf, summed in a fixed
// order) when fewer than min_blocks blocks would run.  Returns -1 on shapes
// the kernel does not cover (nothing is launched then).
bool gemm_shape_ok(int cfg, int M, int N, int K, int lda, int ldb);
int gemm_nt(hipStream_t st, int cfg, const bf16 *A, int lda, const bf16 *B,
            int ldb, float *C, int ldc, int M, int N, int K, float *splitbuf,
            size_t splitbuf_floats, int min_blocks);

// out / outT (BF16, optional) = op(Z) and its transpose; outF (FP32,
// optional) = op(Z).  Z is M x N row-major.
void epilogue(hipStream_t st, const float *Z, int M, int N, int mode,
              const bf16 *mask, bf16 *out, bf16 *outT, float *outF);

// Encoder Conv1d(1 -> C, width T, stride S) + ReLU on X [M x nk] (FP32):
// H0 [M x P0*C] and H0T [P0*C x M] (optional), feature p*C + c.
void conv_forward(hipStream_t st, const float *X, int M, int nk, const float *Wc,
                  int C, int T, int S, int P0, bf16 *H0, bf16 *H0T);
void conv_backward(hipStream_t st, const float *X, int M, int nk, const float *dZ0,
                   int C, int T, int S, int P0, float *gWc);

// Mean squared error over rows < B and columns < nout of Y - Tg (row pitch
// ld); Y becomes dLoss/dY; *loss_out (device) = the mean.
void mse(hipStream_t st, float *Y, const float *Tg, int Mb, int B, int nout,
         int ld, double *partial, double *loss_out);

// Rows b < B of X [Mb x nk] and Tg [Mb x ld] from pair idx[b] of the store;
// other rows zero.
void gather(hipStream_t st, const float *xs, const float *ts, const int *idx,
            int B, int Mb, int nk, int ld, float *X, float *Tg);

void adam(hipStream_t st, float *w, float *m, float *v, const float *g,
          long long n, float lr, float b1, float b2, float eps, float c1,
          float c2);

}  // namespace nn
}  // namespace qdeq
