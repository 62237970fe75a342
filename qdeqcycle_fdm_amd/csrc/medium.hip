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
#include <mutex>
#include <vector>

#include "qdeq_internal.h"

namespace qdeq {

constexpr int NSAMP = 400;

// Sample positions of the interpolation nodes: 3 nodes at -1, 0, 1 or 4 at
// -1.5, -0.5, 0.5, 1.5 (odeSolver.cpp:423-424, 466-467), exact in binary.
__host__ __device__ inline double node_pos(int n, int i) {
  return n == 3 ? (double)(i - 1) : (double)i - 1.5;
}

// Weights w[i*NJ + j] = Lx[i] * Ly[j] of sample s (0..399): mj outer, mi
// inner, both in -10..10 without 0, as interp_2d forms them
// (odeSolver.cpp:134-200).  Host and device evaluate the same operations in
// the same order, so the two tables hold the same bits.
template <int NI, int NJ>
__host__ __device__ inline void lagrange_sample(int s, double *w) {
  const int NG = 10;
  const int a = s / (2 * NG), c = s - a * (2 * NG);
  const int mj = a < NG ? a - NG : a - NG + 1;
  const int mi = c < NG ? c - NG : c - NG + 1;
  const double xi = mi / ((NG + 1.0) * 2.0);
  const double yi = mj / ((NG + 1.0) * 2.0);
  double Lx[NI], Ly[NJ], t[4], b[4];
  for (int i = 0; i < NI; i++) {
    Lx[i] = 1.0;
    for (int k = 0; k < NI; k++) b[k] = node_pos(NI, i) - node_pos(NI, k);
    b[i] = 1.0;
    for (int k = 0; k < NI; k++) t[k] = xi - node_pos(NI, k);
    t[i] = 1.0;
    for (int k = 0; k < NI; k++) Lx[i] = Lx[i] * t[k] / b[k];
  }
  for (int i = 0; i < NJ; i++) {
    Ly[i] = 1.0;
    for (int k = 0; k < NJ; k++) b[k] = node_pos(NJ, i) - node_pos(NJ, k);
    b[i] = 1.0;
    for (int k = 0; k < NJ; k++) t[k] = yi - node_pos(NJ, k);
    t[i] = 1.0;
    for (int k = 0; k < NJ; k++) Ly[i] = Ly[i] * t[k] / b[k];
  }
  for (int i = 0; i < NI; i++)
    for (int j = 0; j < NJ; j++) w[i * NJ + j] = Lx[i] * Ly[j];
}

// Both tables: x-faces (3 x 4 nodes) then z-faces (4 x 3 nodes), 400 x 12
// doubles each.
constexpr int kTableWords = NSAMP * 12;

static void lagrange_tables_host(std::vector<double> &w) {
  w.assign((size_t)2 * kTableWords, 0.0);
  for (int s = 0; s < NSAMP; s++) {
    lagrange_sample<3, 4>(s, &w[(size_t)s * 12]);
    lagrange_sample<4, 3>(s, &w[(size_t)kTableWords + (size_t)s * 12]);
  }
}

__global__ void k_lagrange_tables(double *__restrict__ w) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= NSAMP) return;
  lagrange_sample<3, 4>(s, w + (size_t)s * 12);
  lagrange_sample<4, 3>(s, w + (size_t)kTableWords + (size_t)s * 12);
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

// mh: nodal mu (nz, nx); only rows 1..nk, columns 1..ni are read, each ghost
// taking the value of the nearest interior node (the ghost fill of
// odeSolver.cpp:347-376).  me: doubled array.
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
#define MH(row, col) \
  mh[(size_t)min(max((row), 1), nk) * nx + min(max((col), 1), ni)]
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

// Nodal entries of the doubled array, ghosts included (odeSolver.cpp:337-376):
// a ghost takes the value of the nearest interior node.
__global__ void k_nodal(const double *__restrict__ mh, double *__restrict__ me,
                        int nx, int nz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (i >= nx || k >= nz) return;
  const int ic = min(max(i, 1), nx - 2), kc = min(max(k, 1), nz - 2);
  me[2L * i + 2L * k * (2L * nx - 1)] = mh[(size_t)kc * nx + ic];
}

// The launches of the build: zeroed output, nodal entries, the four face
// kinds.  mh: nodal mu on the device; w: both weight tables on the device.
static void launch_mu_equi(int ni, int nk, const double *mh, double *me,
                           const double *w, hipStream_t s) {
  const int nx = ni + 2, nz = nk + 2;
  const dim3 blk(64);
  const unsigned gx = (unsigned)((ni + 63) / 64);
  hipLaunchKernelGGL(k_nodal, dim3((unsigned)((nx + 255) / 256), nz), dim3(256),
                     0, s, mh, me, nx, nz);
  hipLaunchKernelGGL(k_faces, dim3(gx, nk), blk, 0, s, mh, me, w, w + kTableWords,
                     ni, nk, 0);
  hipLaunchKernelGGL(k_faces, dim3(gx, nk - 1), blk, 0, s, mh, me, w,
                     w + kTableWords, ni, nk, 1);
  hipLaunchKernelGGL(k_faces, dim3(gx, 1), blk, 0, s, mh, me, w, w + kTableWords,
                     ni, nk, 2);
  hipLaunchKernelGGL(k_faces, dim3(gx, 1), blk, 0, s, mh, me, w, w + kTableWords,
                     ni, nk, 3);
}

// mu_host: nodal (nz, nx) mu on the host (ghost entries ignored).
int build_mu_equi_device(int ni, int nk, const double *mu_host, double *me_dev,
                         hipStream_t s, std::string *err) {
  const int nx = ni + 2, nz = nk + 2;
  std::vector<double> w;
  lagrange_tables_host(w);
  double *d_mh = nullptr, *d_w = nullptr;
  const size_t nme = (size_t)(2 * nx - 1) * (2 * nz - 1);
  const size_t nm = (size_t)nx * nz;
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
  MQ_CHECK(hipMalloc(&d_mh, nm * sizeof(double)));
  MQ_CHECK(hipMalloc(&d_w, w.size() * sizeof(double)));
  MQ_CHECK(hipMemcpyAsync(d_mh, mu_host, nm * sizeof(double), hipMemcpyHostToDevice, s));
  MQ_CHECK(hipMemcpyAsync(d_w, w.data(), w.size() * sizeof(double),
                          hipMemcpyHostToDevice, s));
  MQ_CHECK(hipMemsetAsync(me_dev, 0, nme * sizeof(double), s));
  launch_mu_equi(ni, nk, d_mh, me_dev, d_w, s);
  MQ_CHECK(hipGetLastError());
  MQ_CHECK(hipStreamSynchronize(s));
  (void)hipFree(d_mh);
  (void)hipFree(d_w);
#undef MQ_CHECK
  return QDEQ_OK;
}

// ---- the build without allocation ----
//
// The two weight tables live in a file-scope device array: each device the
// library runs on holds its own copy until the process ends.  Host code
// never uses the array's name as a device address; it looks the address up
// once per device (hipGetSymbolAddress, with that device current) and passes
// that.  Every call recomputes the tables into it, so nothing depends on an
// earlier call; calls that overlap on one device share it and race.
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[2 * kTableWords];
static double *g_scratch[kMaxDevices];
static std::mutex g_scratch_mutex;

static int scratch_on(int device, double **out, std::string *err) {
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
      *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return QDEQ_ERR_HIP;
    }
    g_scratch[device] = static_cast<double *>(p);
  }
  *out = g_scratch[device];
  return QDEQ_OK;
}

int build_mu_equi_noalloc(int ni, int nk, const double *mu_dev, double *me_dev,
                          int device, std::string *err) {
  if (device < 0 || device >= kMaxDevices) {
    *err = "device ordinal out of range";
    return QDEQ_ERR_ARG;
  }
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) {
    *err = hipGetErrorString(e);
    return QDEQ_ERR_HIP;
  }
  double *w = nullptr;
  int st = scratch_on(device, &w, err);
  if (st != QDEQ_OK) return st;
  const size_t nme = (size_t)(2 * (ni + 2) - 1) * (2 * (nk + 2) - 1);
  const hipStream_t s = nullptr;  // the null stream: no stream is created
  hipLaunchKernelGGL(k_lagrange_tables, dim3((NSAMP + 127) / 128), dim3(128), 0, s, w);
  e = hipMemsetAsync(me_dev, 0, nme * sizeof(double), s);
  if (e == hipSuccess) {
    launch_mu_equi(ni, nk, mu_dev, me_dev, w, s);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    *err = hipGetErrorString(e);
    return QDEQ_ERR_HIP;
  }
  return QDEQ_OK;
}

}  // namespace qdeq
