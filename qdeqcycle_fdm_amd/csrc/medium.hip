// medium.hip -- the equivalent-medium coefficient array mu_equi on the GPU.
//
// XuSun-SUSTech/QDEqCycle_FDM SetMedium (odeSolver.cpp:337-589) assigns each
// cell face the harmonic mean of 400 samples of a tensor Lagrange
// interpolant (odeSolver.cpp:134-268) of the surrounding nodal mu (3x4 nodes
// for x-faces, 4x3 for z-faces).  The interpolation weights depend only on
// the sample position, so they are formed once on the host (same products,
// same order) and every face becomes 400 x 12 multiply-adds: one thread per
// face.  The x-face block entry the reference leaves unassigned
// (odeSolver.cpp:434-436 skip mu_matrix[1][1]) takes the nodal value at that
// position.
#include <vector>

#include "qdeq_internal.h"

namespace qdeq {

constexpr int NSAMP = 400;

// Weights w[s][i*nj + j] = Lx[i] * Ly[j] for sample s (mj outer, mi inner,
// both in -10..10 without 0), as interp_2d forms them (odeSolver.cpp:134-200).
static void lagrange_table(const double *xn, const double *yn, int ni, int nj,
                           std::vector<double> &w) {
  w.assign((size_t)NSAMP * ni * nj, 0.0);
  const int NG = 10;
  int s = 0;
  for (int mj = -NG; mj <= NG; mj++) {
    if (mj == 0) continue;
    for (int mi = -NG; mi <= NG; mi++) {
      if (mi == 0) continue;
      const double xi = mi / ((NG + 1.0) * 2.0);
      const double yi = mj / ((NG + 1.0) * 2.0);
      double Lx[4], Ly[4], t[4], b[4];
      for (int i = 0; i < ni; i++) {
        Lx[i] = 1.0;
        for (int k = 0; k < ni; k++) b[k] = xn[i] - xn[k];
        b[i] = 1.0;
        for (int k = 0; k < ni; k++) t[k] = xi - xn[k];
        t[i] = 1.0;
        for (int k = 0; k < ni; k++) Lx[i] = Lx[i] * t[k] / b[k];
      }
      for (int i = 0; i < nj; i++) {
        Ly[i] = 1.0;
        for (int k = 0; k < nj; k++) b[k] = yn[i] - yn[k];
        b[i] = 1.0;
        for (int k = 0; k < nj; k++) t[k] = yi - yn[k];
        t[i] = 1.0;
        for (int k = 0; k < nj; k++) Ly[i] = Ly[i] * t[k] / b[k];
      }
      for (int i = 0; i < ni; i++)
        for (int j = 0; j < nj; j++) w[(size_t)s * ni * nj + i * nj + j] = Lx[i] * Ly[j];
      s++;
    }
  }
}

// Harmonic mean over the samples of sum_j sum_i w z (odeSolver.cpp:442-460).
template <int NI, int NJ>
__device__ inline double face_mean(const double *__restrict__ w,
                                   const double (&z)[NI * NJ]) {
  double mu0 = 0.0;
  for (int s = 0; s < NSAMP; s++) {
    const double *ws = w + s * NI * NJ;
    double zi = 0.0;
#pragma unroll
    for (int j = 0; j < NJ; j++)
#pragma unroll
      for (int i = 0; i < NI; i++) zi = zi + ws[i * NJ + j] * z[i * NJ + j];
    mu0 = mu0 + 1.0 / zi;
  }
  return NSAMP / mu0;
}

// mh: nodal mu with the ghost layer filled, (nz, nx).  me: doubled array.
// kind 0: x-faces between nodes (i-1, i), rows 1..nk, i = 2..ni
//      (odeSolver.cpp:421-462); kind 1: z-faces between rows (k-1, k),
//      k = 2..nk, i = 2..ni-1 (464-506); kind 2/3: bottom/top boundary
//      z-faces (509-548, 550-589).
__global__ void k_faces(const double *__restrict__ mh, double *__restrict__ me,
                        const double *__restrict__ wx,
                        const double *__restrict__ wz, int ni, int nk,
                        int kind) {
  const int nx = ni + 2;
  const long W2 = 2L * nx - 1;
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 2;
  const int r = blockIdx.y;
#define MH(row, col) mh[(size_t)(row) * nx + (col)]
  if (kind == 0) {
    const int k = r + 1;
    if (i > ni || k > nk) return;
    double z[12];
    for (int q = 0; q < 4; q++) {
      z[0 * 4 + q] = MH(k - 1, i - 2 + q);
      z[1 * 4 + q] = MH(k, i - 2 + q);
      z[2 * 4 + q] = MH(k + 1, i - 2 + q);
    }
    me[(2L * i - 1) + (2L * k) * W2] = face_mean<3, 4>(wx, z);
    return;
  }
  if (i > ni - 1) return;
  int r0, r1, r2, r3;
  long index;
  if (kind == 1) {
    const int k = r + 2;
    if (k > nk) return;
    r0 = k - 2; r1 = k - 1; r2 = k; r3 = k + 1;
    index = 2L * i + (2L * k - 1) * W2;
  } else if (kind == 2) {
    r0 = 0; r1 = 0; r2 = 1; r3 = 2;
    index = 2L * i + 1 * W2;
  } else {
    const int nk2 = nk + 1;
    r0 = nk2 - 2; r1 = nk2 - 1; r2 = nk2; r3 = nk2;
    index = 2L * i + (2L * nk2 - 1) * W2;
  }
  double z[12];
  const int rr[4] = {r0, r1, r2, r3};
  for (int p = 0; p < 4; p++)
    for (int q = 0; q < 3; q++) z[p * 3 + q] = MH(rr[p], i - 1 + q);
  me[index] = face_mean<4, 3>(wz, z);
#undef MH
}

// Nodal entries of the doubled array, ghosts included (odeSolver.cpp:337-376).
__global__ void k_nodal(const double *__restrict__ mh, double *__restrict__ me,
                        int nx, int nz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (i >= nx || k >= nz) return;
  me[2L * i + 2L * k * (2L * nx - 1)] = mh[(size_t)k * nx + i];
}

// mu_dev: nodal (nz, nx) mu on the device (ghost entries ignored).
int build_mu_equi_device(int ni, int nk, const double *mu_host, double *me_dev,
                         hipStream_t s, std::string *err) {
  const int nx = ni + 2, nz = nk + 2;
  // ghost fill of the nodal values (odeSolver.cpp:347-376)
  std::vector<double> mh((size_t)nx * nz, 0.0);
  for (int k = 1; k <= nk; k++)
    for (int i = 1; i <= ni; i++) mh[(size_t)k * nx + i] = mu_host[(size_t)k * nx + i];
  for (int i = 1; i <= ni; i++) {
    mh[i] = mh[(size_t)1 * nx + i];
    mh[(size_t)(nz - 1) * nx + i] = mh[(size_t)nk * nx + i];
  }
  for (int k = 0; k < nz; k++) {
    mh[(size_t)k * nx] = mh[(size_t)k * nx + 1];
    mh[(size_t)k * nx + nx - 1] = mh[(size_t)k * nx + ni];
  }
  static const double x_index[3] = {-1.0, 0.0, 1.0};
  static const double y_index[4] = {-1.5, -0.5, 0.5, 1.5};
  static const double x_index2[4] = {-1.5, -0.5, 0.5, 1.5};
  static const double y_index2[3] = {-1.0, 0.0, 1.0};
  std::vector<double> wx, wz;
  lagrange_table(x_index, y_index, 3, 4, wx);
  lagrange_table(x_index2, y_index2, 4, 3, wz);

  double *d_mh = nullptr, *d_w = nullptr;
  const size_t nme = (size_t)(2 * nx - 1) * (2 * nz - 1);
  hipError_t e;
#define MQ_CHECK(call)                                               \
  do {                                                               \
    e = (call);                                                      \
    if (e != hipSuccess) {                                           \
      if (err) *err = std::string(#call) + ": " + hipGetErrorString(e); \
      (void)hipFree(d_mh);                                            \
      (void)hipFree(d_w);                                            \
      return QDEQ_ERR_HIP;                                           \
    }                                                                \
  } while (0)
  MQ_CHECK(hipMalloc(&d_mh, mh.size() * sizeof(double)));
  MQ_CHECK(hipMalloc(&d_w, (wx.size() + wz.size()) * sizeof(double)));
  MQ_CHECK(hipMemcpyAsync(d_mh, mh.data(), mh.size() * sizeof(double),
                          hipMemcpyHostToDevice, s));
  MQ_CHECK(hipMemcpyAsync(d_w, wx.data(), wx.size() * sizeof(double),
                          hipMemcpyHostToDevice, s));
  MQ_CHECK(hipMemcpyAsync(d_w + wx.size(), wz.data(), wz.size() * sizeof(double),
                          hipMemcpyHostToDevice, s));
  MQ_CHECK(hipMemsetAsync(me_dev, 0, nme * sizeof(double), s));
  hipLaunchKernelGGL(k_nodal, dim3((unsigned)((nx + 255) / 256), nz), dim3(256),
                     0, s, d_mh, me_dev, nx, nz);
  const dim3 blk(64);
  const unsigned gx = (unsigned)((ni + 63) / 64);
  hipLaunchKernelGGL(k_faces, dim3(gx, nk), blk, 0, s, d_mh, me_dev, d_w,
                     d_w + wx.size(), ni, nk, 0);
  hipLaunchKernelGGL(k_faces, dim3(gx, nk - 1), blk, 0, s, d_mh, me_dev, d_w,
                     d_w + wx.size(), ni, nk, 1);
  hipLaunchKernelGGL(k_faces, dim3(gx, 1), blk, 0, s, d_mh, me_dev, d_w,
                     d_w + wx.size(), ni, nk, 2);
  hipLaunchKernelGGL(k_faces, dim3(gx, 1), blk, 0, s, d_mh, me_dev, d_w,
                     d_w + wx.size(), ni, nk, 3);
  MQ_CHECK(hipGetLastError());
  MQ_CHECK(hipStreamSynchronize(s));
  (void)hipFree(d_mh);
  (void)hipFree(d_w);
#undef MQ_CHECK
  return QDEQ_OK;
}

}  // namespace qdeq
