// coarse.hip -- exact solve on the coarsest level of the fast cycle.
//
// With n_coarse == 0 the fast cycle replaces the coarsest-level sweeps by a
// direct solve: the coarsest operator (the same 5-point stencil the sweeps
// use, Dirichlet columns removed, Neumann rows mirrored) is assembled on the
// host, inverted once by Gauss-Jordan elimination, and applied as a dense
// matrix-vector product.  The inverse is stored in FP32: the product is
// memory-bound in one workgroup (255^2 entries at 1025^2 with 7 levels), and
// a coarse correction accurate to ~1e-7 relative leaves the V-cycle's
// contraction unchanged; the FP64 solve multiplies and sums in FP64.  Each
// row of the product is summed in column order with separate multiplies and
// adds, so the result is reproducible bit for bit by a host implementation
// of the same steps.
#include <vector>

#include "qdeq_internal.h"

namespace qdeq {

// B = DH^2 * L on the unknowns (j, k), k in [1, K-2], index j*(K-2) + k-1.
static void assemble_coarse(const std::vector<double> &cx,
                            const std::vector<double> &cz, int N, int K, int P,
                            std::vector<double> &B) {
  const int W = K - 2, n = N * W;
  B.assign((size_t)n * n, 0.0);
  for (int j = 0; j < N; j++)
    for (int k = 1; k <= K - 2; k++) {
      const int row = j * W + (k - 1);
      const size_t g = (size_t)j * P + k;
      const double m1 = cx[g], m2 = cx[g + 1], m3 = cz[g], m4 = cz[g + P];
      const int jS = (j == 0) ? 1 : j - 1;
      const int jN = (j == N - 1) ? N - 2 : j + 1;
      double *Br = &B[(size_t)row * n];
      if (k - 1 >= 1) Br[j * W + (k - 2)] += m1;
      if (k + 1 <= K - 2) Br[j * W + k] += m2;
      Br[jS * W + (k - 1)] += m3;
      Br[jN * W + (k - 1)] += m4;
      Br[row] -= (m1 + m2 + m3 + m4);
    }
}

// In-place Gauss-Jordan inverse without pivoting (the operator is a
// diagonally dominant M-matrix); X receives A^{-1}, row-major.
static void gauss_jordan(std::vector<double> &A, int n, std::vector<double> &X) {
  X.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) X[(size_t)i * n + i] = 1.0;
  for (int p = 0; p < n; p++) {
    double *Ap = &A[(size_t)p * n], *Xp = &X[(size_t)p * n];
    const double piv = Ap[p];
    for (int c = 0; c < n; c++) {
      Ap[c] = Ap[c] / piv;
      Xp[c] = Xp[c] / piv;
    }
    for (int r = 0; r < n; r++) {
      if (r == p) continue;
      double *Ar = &A[(size_t)r * n], *Xr = &X[(size_t)r * n];
      const double f = Ar[p];
      if (f == 0.0) continue;
      for (int c = 0; c < n; c++) {
        Ar[c] = Ar[c] - f * Ap[c];
        Xr[c] = Xr[c] - f * Xp[c];
      }
    }
  }
}

int build_coarse_inverse(qdeq_ctx *c) {
  (void)hipFree(c->cinv32);
  c->cinv32 = nullptr;
  c->coarse_n = 0;
  if (c->cfg.cycle != QDEQ_CYCLE_FAST || c->cfg.n_coarse != 0) return QDEQ_OK;
  const LevelDims &d = c->lv[c->L - 1];
  const int n = d.N * (d.K - 2);
  if (n > 1024)
    return set_error(c, QDEQ_ERR_ARG,
                     "exact coarse solve needs at most 1024 coarsest unknowns "
                     "(use more levels or n_coarse > 0)");
  std::vector<double> cx((size_t)d.N * d.P), cz((size_t)(d.N + 1) * d.P);
  QDEQ_HIP(c, hipMemcpyAsync(cx.data(), c->d[c->L - 1].cx, cx.size() * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
  QDEQ_HIP(c, hipMemcpyAsync(cz.data(), c->d[c->L - 1].cz, cz.size() * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
  QDEQ_HIP(c, hipStreamSynchronize(c->stream));
  std::vector<double> B, X;
  assemble_coarse(cx, cz, d.N, d.K, d.P, B);
  gauss_jordan(B, n, X);
  // column-major FP32 copy: element (row, col) at col*n + row
  std::vector<float> xt32((size_t)n * n);
  for (int r = 0; r < n; r++)
    for (int col = 0; col < n; col++)
      xt32[(size_t)col * n + r] = (float)X[(size_t)r * n + col];
  QDEQ_HIP(c, hipMalloc(&c->cinv32, xt32.size() * sizeof(float)));
  QDEQ_HIP(c, hipMemcpyAsync(c->cinv32, xt32.data(), xt32.size() * sizeof(float),
                             hipMemcpyHostToDevice, c->stream));
  QDEQ_HIP(c, hipStreamSynchronize(c->stream));
  c->coarse_n = n;
  return QDEQ_OK;
}

// One-block coarse solve for the untiled cycle: u = B^{-1} (DH^2 f).
template <typename T>
__global__ __launch_bounds__(1024) void k_coarse_direct(
    T *__restrict__ u, const T *__restrict__ f, const float *__restrict__ minvT,
    int N, int K, int P, T DH2, int n) {
  __shared__ T g[1024];
  __shared__ T part[1024];
  const int W = K - 2;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int j = i / W, k = i - j * W + 1;
    g[i] = DH2 * f[(size_t)j * P + k];
  }
  __syncthreads();
  coarse_matvec<T>(minvT, g, part, n, u, W, K, (size_t)P);
}

void launch_coarse_direct_fp64(qdeq_ctx *c, double *u, const double *f) {
  const LevelDims &d = c->lv[c->L - 1];
  hipLaunchKernelGGL(k_coarse_direct<double>, dim3(1), dim3(1024), 0, c->stream,
                     u, f, c->cinv32, d.N, d.K, d.P, d.DH * d.DH, c->coarse_n);
}

void launch_coarse_direct_fp32(qdeq_ctx *c, float *u, const float *f) {
  const LevelDims &d = c->lv[c->L - 1];
  hipLaunchKernelGGL(k_coarse_direct<float>, dim3(1), dim3(1024), 0, c->stream,
                     u, f, c->cinv32, d.N, d.K, d.P, (float)(d.DH * d.DH),
                     c->coarse_n);
}

}  // namespace qdeq
