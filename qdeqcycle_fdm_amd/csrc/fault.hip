// fault.hip -- fault traction, rate-and-state slip rate, state rate and the
// RK43 stage arithmetic, one thread per fault node, FP64.
//
// Follows XuSun-SUSTech/QDEqCycle_FDM fault.cpp:139-232 and
// odeSolver.cpp:1150-1206, 1491-1614.  Expressions keep the reference's
// operation order (-ffp-contract=off); exp/sinh/cosh/asinh are the
// host/device-identical implementations of qd_math.h.
#include <cmath>

#include "qd_math.h"
#include "qdeq_internal.h"

namespace qdeq {

constexpr int FT = 256;  // threads per block for per-node kernels

static inline dim3 nodes_grid(int nk) { return dim3((unsigned)((nk + FT - 1) / FT)); }

// Butcher tableau (odeSolver.cpp:1297-1334).
__constant__ double kA[7][6] = {
    {0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0},
    {0, 1. / 2., 0, 0, 0, 0},
    {0, 13861. / 62500., 6889. / 62500., 0, 0, 0},
    {0, -116923316275. / 2393684061468., -2731218467317. / 15368042101831.,
     9408046702089. / 11113171139209., 0, 0},
    {0, -451086348788. / 2902428689909., -2682348792572. / 7519795681897.,
     12662868775082. / 11960479115383., 3355817975965. / 11060851509271., 0},
    {0, 647845179188. / 3216320057751., 73281519250. / 8382639484533.,
     552539513391. / 3454668386233., 3354512671639. / 8306763924573.,
     4040. / 17871.}};
static const double kC[7] = {0, 0, 1. / 2., 83. / 250., 31. / 50., 17. / 20., 1.};
constexpr double B1 = 82889. / 524892., B3 = 15625. / 83664.,
                 B4 = 69875. / 102672., B5 = -2260. / 8211., B6 = 1. / 4.;
constexpr double HB1 = 4586570599. / 29645900160., HB3 = 178811875. / 945068544.,
                 HB4 = 814220225. / 1159782912., HB5 = -3700637. / 11593932.,
                 HB6 = 61727. / 225920.;

// NRsearch for one node (fault.cpp:180-228).  Returns false on NaN.
__device__ inline bool nr_node(double Ts, double Tn, double psi, double a,
                               double eta, double V0, double Vprev,
                               double *Vout) {
  // e^(-psi/a) and 2 eta V0 e^(-psi/a) of every iteration's two expressions,
  // formed once with their left-to-right products, so the same bits
  const double em = qd_exp(-psi / a);
  const double c = 2.0 * eta * V0 * em;
  for (int attempt = 0; attempt < 2; attempt++) {  // the reference retries once
    int iter = 0;
    double var_old = 0.5 * Vprev / V0 * qd_exp(psi / a);
    var_old = qd_asinh(var_old);
    const double fa = a * fabs(Tn);
    const double fb = Ts;
    double var_new = var_old;
    const int iter_max = 100000;
    while (iter < iter_max) {
      double sh, ch;
      qd_sinh_cosh(var_old, &sh, &ch);
      const double fy = c * sh + fa * var_old - fb;
      const double df = c * ch + fa;
      const double d = -fy / (df + 1e-100);
      var_new = var_old + d;
      const double err = fabs(d) / (fabs(var_old) + 1e-100);
      iter = iter + 1;
      var_old = var_new;
      if (err < 1e-10) break;
    }
    const double friction_update = a * var_new;
    if (!isnan(friction_update)) {
      const double V = 2.0 * V0 * em;
      *Vout = V * qd_sinh(var_new);
      return true;
    }
  }
  return false;
}

// cal_dpsi for one node (odeSolver.cpp:1198-1203).
__device__ inline double dpsi_node(double psi, double V, double b, double f0,
                                   double V0, double L) {
  const double A = qd_exp((f0 - psi) / b);
  double d = 0.0;
  if (!isinf(A) && b > 1e-3) d = (b * V0 / L) * (A - V / V0);
  return d;
}

// ---- drop-in kernels on the reference's nx*nz grid array ----

// setBoundary (odeSolver.cpp:1150-1188): boundary columns; ghost rows after.
__global__ void k_set_boundary_cols(double *U, int ni, int nk,
                                    const double *slip, const double *ss,
                                    double Vini, double t, double ratio,
                                    double dt) {
  const int nx = ni + 2, nz = nk + 2;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nz) return;
  if (k >= 1 && k <= nk) U[(size_t)k * nx + 1] = 0.5 * slip[k];
  U[(size_t)k * nx + ni] = 0.50 * Vini * (t + ratio * dt) + ss[k];
}
__global__ void k_set_boundary_rows(double *U, int ni, int nk) {
  const int nx = ni + 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx) return;
  U[i] = U[(size_t)2 * nx + i];
  U[(size_t)(nk + 1) * nx + i] = U[(size_t)(nk - 1) * nx + i];
}

// extend_crew (odeSolver.cpp:612-650): columns first, then rows.
__global__ void k_extend_cols(double *U, int ni, int nk) {
  const int nx = ni + 2, nz = nk + 2;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nz) return;
  const size_t i1 = 1 + (size_t)k * nx, i2 = ni + (size_t)k * nx;
  U[i1 - 1] = 2.0 * U[i1] - U[i1 + 1];
  U[i2 + 1] = 2.0 * U[i2] - U[i2 - 1];
}
__global__ void k_extend_rows(double *U, int ni, int nk) {
  const int nx = ni + 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx) return;
  U[i] = U[(size_t)2 * nx + i];
  U[(size_t)(nk + 1) * nx + i] = U[(size_t)(nk - 1) * nx + i];
}

// pack/unpack between the nx*nz grid array and the compact level-0 array
// (odeSolver.cpp:669-677, 708-716).
__global__ void k_pack(const double *U, double *u, int ni, int nk, int P) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // compact column
  const int j = blockIdx.y;                              // compact row
  if (i >= ni || j >= nk) return;
  u[(size_t)j * P + i] = U[(size_t)(j + 1) * (ni + 2) + i + 1];
}
__global__ void k_unpack(double *U, const double *u, int ni, int nk, int P) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= ni || j >= nk) return;
  U[(size_t)(j + 1) * (ni + 2) + i + 1] = u[(size_t)j * P + i];
}

// eval_stress_on_fault + partial1st (fault.cpp:139-164) on a grid array.
__global__ void k_eval_stress_grid(const double *U, const double *mu,
                                   const double *tau_n0, int ni, int nk,
                                   double c0, double c1, double c2, double c3,
                                   double *Ts, double *Tn) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (k > nk) return;
  const size_t index = 1 + (size_t)k * (ni + 2);
  const double du =
      c0 * U[index] + c1 * U[index + 1] + c2 * U[index + 2] + c3 * U[index + 3];
  Ts[k] = mu[index] * du * 1.0e3;
  Tn[k] = tau_n0[k];
}

__global__ void k_nr(const double *Ts, const double *Tn, const double *psi,
                     const double *a, const double *eta, double V0, double *V,
                     int nk, int *flags) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (k > nk) return;
  double Vn;
  if (nr_node(Ts[k], Tn[k], psi[k], a[k], eta[k], V0, V[k], &Vn))
    V[k] = Vn;
  else
    flags[0] = 1;
}

__global__ void k_dpsi(const double *psi, const double *V, const double *b,
                       double f0, double V0, double L, double *dpsi, int nk) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (k > nk) return;
  dpsi[k] = dpsi_node(psi[k], V[k], b[k], f0, V0, L);
}

// ---- device-resident RK43 stages ----

// Stage state Y_s = y_n + dt sum_j a_sj F_j (odeSolver.cpp:1491-1561) and
// its boundary data written into the compact level-0 u (setBoundary with
// ratio c_s: fault column slip/2, remote column Vini (t + c_s dt)/2 +
// SlipShift).  s = 1 is the committed state (odeSolver.cpp:1615).
__device__ inline void stage_node(int k, int s, double dt, double t, double ratio,
                                  double Vini, const double *slip,
                                  const double *psi, const double *ss,
                                  double *const *FV, double *const *Fd,
                                  double *Ys, double *Yp, double *u, int K,
                                  int P) {
  double ys = slip[k], yp = psi[k];
  if (s >= 2) {
    ys = slip[k] + kA[s][1] * dt * FV[1][k];
    yp = psi[k] + kA[s][1] * dt * Fd[1][k];
    for (int jj = 2; jj < s; jj++) {
      ys = ys + kA[s][jj] * dt * FV[jj][k];
      yp = yp + kA[s][jj] * dt * Fd[jj][k];
    }
  }
  Ys[k] = ys;
  Yp[k] = yp;
  double *row = u + (size_t)(k - 1) * P;
  row[0] = 0.5 * ys;
  row[K - 1] = 0.50 * Vini * (t + ratio * dt) + ss[k];
}

// stage_node with its inputs in registers: slip, psi and SlipShift of the
// node and the stage rates F_1 .. F_{s-1} (fv[j], fd[j]); the same
// operations in the same order, so the same bits.
__device__ inline void stage_node_in(int k, int s, double dt, double t, double ratio,
                                     double Vini, double sl, double ps, double ssk,
                                     const double (&fv)[6], const double (&fd)[6],
                                     double *Ys, double *Yp, double *u, int K, int P) {
  double ys = sl, yp = ps;
  if (s >= 2) {
    ys = sl + kA[s][1] * dt * fv[1];
    yp = ps + kA[s][1] * dt * fd[1];
#pragma unroll
    for (int jj = 2; jj < 6; jj++) {
      if (jj < s) {
        ys = ys + kA[s][jj] * dt * fv[jj];
        yp = yp + kA[s][jj] * dt * fd[jj];
      }
    }
  }
  Ys[k] = ys;
  Yp[k] = yp;
  double *row = u + (size_t)(k - 1) * P;
  row[0] = 0.5 * ys;
  row[K - 1] = 0.50 * Vini * (t + ratio * dt) + ssk;
}

// tdt: t and dt of the attempt on the device
// go: launched ahead of the accept decision, it runs only if the attempt was
// accepted on the device (go[7] == 1, written by k_combine)
__global__ void k_stage(int s, const double *tdt, double ratio, double Vini,
                        const double *slip, const double *psi,
                        const double *ss, double *const *FV,
                        double *const *Fd, double *Ys, double *Yp, double *u,
                        int nk, int K, int P, const int *go) {
  if (go && go[7] != 1) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (k > nk) return;
  stage_node(k, s, tdt[1], tdt[0], ratio, Vini, slip, psi, ss, FV, Fd, Ys, Yp, u, K,
             P);
}

// Fault update after a stage solve: Ts from the compact u (columns 0..3 are
// the fault and its three neighbours), NR for V from the stage's previous
// rate (zeroed, odeSolver.cpp:1484-1488), then dpsi.
// Returns the node's new rate and state rate (also stored).
struct NodeRates {
  double V, dpsi;
};
__device__ inline NodeRates update_node(int k, const double *u, int P,
                                        const double *mu_f, const double *tau_n0,
                                        const double *Yp, const double *a,
                                        const double *b, const double *eta, double V0,
                                        double f0, double L, double c0, double c1,
                                        double c2, double c3, double *Ts, double *Tn,
                                        double *V, double *dpsi, int *flags) {
  const double *row = u + (size_t)(k - 1) * P;
  const double du = c0 * row[0] + c1 * row[1] + c2 * row[2] + c3 * row[3];
  const double ts = mu_f[k] * du * 1.0e3;
  const double tn = tau_n0[k];
  Ts[k] = ts;
  Tn[k] = tn;
  const double psi = Yp[k];
  double Vn;
  if (!nr_node(ts, tn, psi, a[k], eta[k], V0, V[k], &Vn)) {
    flags[0] = 1;
    Vn = V[k];
  }
  V[k] = Vn;
  const double dp = dpsi_node(psi, Vn, b[k], f0, V0, L);
  dpsi[k] = dp;
  return NodeRates{Vn, dp};
}

__global__ void k_fault_update(const double *u, int P, int nk,
                               const double *mu_f, const double *tau_n0,
                               const double *Yp, const double *a,
                               const double *b, const double *eta, double V0,
                               double f0, double L, double c0, double c1,
                               double c2, double c3, double *Ts, double *Tn,
                               double *V, double *dpsi, int *flags) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (k > nk) return;
  update_node(k, u, P, mu_f, tau_n0, Yp, a, b, eta, V0, f0, L, c0, c1, c2, c3, Ts,
              Tn, V, dpsi, flags);
}

// The fault update after the solve of stage s and the stage state and
// boundary data of stage s + 1, in one launch: both are per node (node k
// reads and writes only its own row of u and its own entries), so each
// thread runs the two in order.
struct UpdateArgs {
  const double *mu_f, *tau_n0, *Yp, *a, *b, *eta;
  double V0, f0, L, c0, c1, c2, c3;
  double *Ts, *Tn, *V, *dpsi;
  int *flags;
};
struct StageArgs {
  int s;
  const double *tdt;  // t, dt of the attempt
  double ratio, Vini;
  const double *slip, *psi, *ss;
  double *const *FV;
  double *const *Fd;
  double *Ys, *Yp;
};

__global__ void k_update_stage(double *u, int nk, int K, int P, UpdateArgs ua,
                               StageArgs sa, const int *spec) {
  // launched ahead, behind a solve still in flight: runs only if that solve
  // ended converged (its outcome in spec[2])
  if (spec && spec[2] != 1) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (k > nk) return;
  // the next stage's inputs that the update leaves unchanged, loaded before
  // it (after its stores the loads could not be issued early: the pointers
  // may alias); the rates of the stage just solved (index s - 1 of FV and
  // Fd, the update's V and dpsi) come from the update itself
  const int s = sa.s;  // 2 .. 6
  const double t = sa.tdt[0], dt = sa.tdt[1];
  const double sl = sa.slip[k], ps = sa.psi[k], ssk = sa.ss[k];
  double fv[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double fd[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int jj = 1; jj < 5; jj++) {
    if (jj < s - 1) {
      fv[jj] = sa.FV[jj][k];
      fd[jj] = sa.Fd[jj][k];
    }
  }
  const NodeRates r = update_node(k, u, P, ua.mu_f, ua.tau_n0, ua.Yp, ua.a, ua.b, ua.eta,
                                  ua.V0, ua.f0, ua.L, ua.c0, ua.c1, ua.c2, ua.c3, ua.Ts,
                                  ua.Tn, ua.V, ua.dpsi, ua.flags);
#pragma unroll
  for (int jj = 1; jj < 6; jj++) {
    if (jj == s - 1) {
      fv[jj] = r.V;
      fd[jj] = r.dpsi;
    }
  }
  stage_node_in(k, s, dt, t, sa.ratio, sa.Vini, sl, ps, ssk, fv, fd, sa.Ys, sa.Yp, u, K,
                P);
}

// The fault update after an accepted step's final stage-1 solve, then the
// next attempt's zeroed stage rates and its stage 2 state and boundary data,
// queued behind that solve: it runs only if the solve ended converged
// (spec[2]).  The attempt's t and dt come by value (thread 0 also stores
// them for the later stages), so no thread reads what another one writes.
__global__ void k_update_next(double *u, int nk, int nz, int K, int P, UpdateArgs ua,
                              StageArgs sa, double t, double dt, double *tdt,
                              const int *spec) {
  if (spec[2] != 1) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k == 0) {
    tdt[0] = t;
    tdt[1] = dt;
  }
  if (k >= nz) return;
  for (int s = 2; s <= 6; s++) {
    sa.FV[s][k] = 0.0;
    sa.Fd[s][k] = 0.0;
  }
  if (k < 1 || k > nk) return;
  // stage 2 reads slip, psi, SlipShift (loaded before the update, as in
  // k_update_stage) and F_1, the rates the update forms
  const double sl = sa.slip[k], ps = sa.psi[k], ssk = sa.ss[k];
  const NodeRates r = update_node(k, u, P, ua.mu_f, ua.tau_n0, ua.Yp, ua.a, ua.b, ua.eta,
                                  ua.V0, ua.f0, ua.L, ua.c0, ua.c1, ua.c2, ua.c3, ua.Ts,
                                  ua.Tn, ua.V, ua.dpsi, ua.flags);
  double fv[6] = {0.0, r.V, 0.0, 0.0, 0.0, 0.0};
  double fd[6] = {0.0, r.dpsi, 0.0, 0.0, 0.0, 0.0};
  stage_node_in(k, 2, dt, t, sa.ratio, sa.Vini, sl, ps, ssk, fv, fd, sa.Ys, sa.Yp, u, K,
                P);
}

__global__ void k_zero_rates(double *const *FV, double *const *Fd, int nz,
                             double *tdt, double t, double dt) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k == 0) {
    tdt[0] = t;
    tdt[1] = dt;
  }
  if (k >= nz) return;
  for (int s = 2; s <= 6; s++) {
    FV[s][k] = 0.0;
    Fd[s][k] = 0.0;
  }
}

// y3 (embedded, hb weights) and y4 (b weights, propagated) and the error
// norm (odeSolver.cpp:1569-1595, computeError 1211-1238).  One block; the
// squared differences are summed by one lane in node order, so totErr has
// the reference's summation order.
__global__ __launch_bounds__(1024) void k_combine(
    const double *tdt, const double *slip, const double *psi, double *const *FV,
    double *const *Fd, double *y3s, double *y3p, double *y4s, double *y4p,
    int nk, double *out, int *gate, double totTol) {
  __shared__ double s1[1024], s2[1024];
  const int tid = threadIdx.x;
  const double dt = tdt[1];
  double e1 = 0.0, e2 = 0.0;  // meaningful in lane 0 only
  for (int base = 1; base <= nk; base += 1024) {
    const int k = base + tid;
    if (k <= nk) {
      double a3 = slip[k] + HB1 * dt * FV[1][k];
      a3 = a3 + HB3 * dt * FV[3][k];
      a3 = a3 + HB4 * dt * FV[4][k];
      a3 = a3 + HB5 * dt * FV[5][k];
      a3 = a3 + HB6 * dt * FV[6][k];
      double p3 = psi[k] + HB1 * dt * Fd[1][k];
      p3 = p3 + HB3 * dt * Fd[3][k];
      p3 = p3 + HB4 * dt * Fd[4][k];
      p3 = p3 + HB5 * dt * Fd[5][k];
      p3 = p3 + HB6 * dt * Fd[6][k];
      double a4 = slip[k] + B1 * dt * FV[1][k];
      a4 = a4 + B3 * dt * FV[3][k];
      a4 = a4 + B4 * dt * FV[4][k];
      a4 = a4 + B5 * dt * FV[5][k];
      a4 = a4 + B6 * dt * FV[6][k];
      double p4 = psi[k] + B1 * dt * Fd[1][k];
      p4 = p4 + B3 * dt * Fd[3][k];
      p4 = p4 + B4 * dt * Fd[4][k];
      p4 = p4 + B5 * dt * Fd[5][k];
      p4 = p4 + B6 * dt * Fd[6][k];
      y3s[k] = a3; y3p[k] = p3; y4s[k] = a4; y4p[k] = p4;
      const double t1 = a3 - a4, t2 = p3 - p4;
      s1[tid] = t1 * t1;
      s2[tid] = t2 * t2;
    }
    __syncthreads();
    if (tid == 0) {
      // the sums in node order; unrolled so the LDS loads issue ahead of
      // the dependent additions
      const int n = (nk - base + 1) < 1024 ? (nk - base + 1) : 1024;
#pragma unroll 16
      for (int i = 0; i < n; i++) {
        e1 = e1 + s1[i];
        e2 = e2 + s2[i];
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    const double r1 = sqrt(e1), r2 = sqrt(e2);
    const double err = r1 / (sqrt((double)nk) * 1.0) + r2 / (sqrt((double)nk) * 1.0);
    out[0] = err;
    // the accept test the host applies first (odeSolver.cpp:1596), for
    // work launched ahead of the host's decision
    gate[7] = err < totTol ? 1 : 0;
  }
}

// Accept: y_n <- y4, F_1 <- 0 (odeSolver.cpp:1606-1614); the device time
// becomes t_new = t + dt.
__global__ void k_commit(double *slip, double *psi, const double *y4s,
                         const double *y4p, double *V1, double *d1, int nk,
                         double *tdt, double t_new, const int *go) {
  if (go && go[7] != 1) return;  // launched ahead: only if accepted
  const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (k == 1) tdt[0] = t_new;
  if (k > nk) return;
  slip[k] = y4s[k];
  psi[k] = y4p[k];
  V1[k] = 0.0;
  d1[k] = 0.0;
}

// maxval (fault.cpp:22-52) of V and Ts.  One block.
__global__ __launch_bounds__(1024) void k_maxima(const double *V,
                                                 const double *Ts, int nk,
                                                 double *out) {
  __shared__ double s1[1024], s2[1024];
  double m1 = V[1], m2 = Ts[1];
  for (int k = threadIdx.x + 1; k <= nk; k += 1024) {
    m1 = fmax(m1, V[k]);
    m2 = fmax(m2, Ts[k]);
  }
  const int tid = threadIdx.x;
  s1[tid] = m1;
  s2[tid] = m2;
  __syncthreads();
  for (int st = 512; st > 0; st >>= 1) {
    if (tid < st) {
      s1[tid] = fmax(s1[tid], s1[tid + st]);
      s2[tid] = fmax(s2[tid], s2[tid + st]);
    }
    __syncthreads();
  }
  if (tid == 0) {
    out[0] = s1[0];
    out[1] = s2[0];
  }
}

// ---- launchers ----

void launch_set_boundary_grid(qdeq_ctx *c, double *U, const double *slip,
                              const double *ss, double t, double ratio,
                              double dt) {
  const int ni = c->cfg.ni, nk = c->cfg.nk;
  hipLaunchKernelGGL(k_set_boundary_cols, nodes_grid(nk + 2), dim3(FT), 0,
                     c->stream, U, ni, nk, slip, ss, c->cfg.Vini, t, ratio, dt);
  hipLaunchKernelGGL(k_set_boundary_rows, nodes_grid(ni + 2), dim3(FT), 0,
                     c->stream, U, ni, nk);
}

void launch_extend_crew(qdeq_ctx *c, double *U) {
  const int ni = c->cfg.ni, nk = c->cfg.nk;
  hipLaunchKernelGGL(k_extend_cols, nodes_grid(nk + 2), dim3(FT), 0, c->stream,
                     U, ni, nk);
  hipLaunchKernelGGL(k_extend_rows, nodes_grid(ni + 2), dim3(FT), 0, c->stream,
                     U, ni, nk);
}

void launch_pack(qdeq_ctx *c, const double *U) {
  hipLaunchKernelGGL(k_pack, dim3((unsigned)((c->cfg.ni + 255) / 256), c->cfg.nk),
                     dim3(256), 0, c->stream, U, c->d[0].u, c->cfg.ni,
                     c->cfg.nk, c->lv[0].P);
}

void launch_unpack(qdeq_ctx *c, double *U) {
  hipLaunchKernelGGL(k_unpack, dim3((unsigned)((c->cfg.ni + 255) / 256), c->cfg.nk),
                     dim3(256), 0, c->stream, U, c->d[0].u, c->cfg.ni,
                     c->cfg.nk, c->lv[0].P);
}

void launch_eval_stress_grid(qdeq_ctx *c, const double *U, const double *mu,
                             const double *tau_n0, double *Ts, double *Tn) {
  hipLaunchKernelGGL(k_eval_stress_grid, nodes_grid(c->cfg.nk), dim3(FT), 0,
                     c->stream, U, mu, tau_n0, c->cfg.ni, c->cfg.nk,
                     c->c_stress[0], c->c_stress[1], c->c_stress[2],
                     c->c_stress[3], Ts, Tn);
}

void launch_nr(qdeq_ctx *c, const double *Ts, const double *Tn,
               const double *psi, const double *a, const double *eta,
               double *V) {
  hipLaunchKernelGGL(k_nr, nodes_grid(c->cfg.nk), dim3(FT), 0, c->stream, Ts,
                     Tn, psi, a, eta, c->cfg.V0, V, c->cfg.nk, c->flags);
}

void launch_dpsi(qdeq_ctx *c, const double *psi, const double *V,
                 const double *b, double *dpsi) {
  hipLaunchKernelGGL(k_dpsi, nodes_grid(c->cfg.nk), dim3(FT), 0, c->stream,
                     psi, V, b, c->cfg.f0, c->cfg.V0, c->cfg.L, dpsi,
                     c->cfg.nk);
}

// Device array-of-pointers for the stage rates lives right after the
// fault buffers (set up in api.hip): fb.base + 40*nz.
static double *const *dev_FV(qdeq_ctx *c) {
  return reinterpret_cast<double *const *>(c->fb.base + (size_t)40 * c->nz);
}
static double *const *dev_Fd(qdeq_ctx *c) { return dev_FV(c) + 7; }

void launch_stage(qdeq_ctx *c, int s, const int *go) {
  FaultBufs &f = c->fb;
  hipLaunchKernelGGL(k_stage, nodes_grid(c->cfg.nk), dim3(FT), 0, c->stream, s,
                     c->scal + SC_T, kC[s], c->cfg.Vini, f.slip, f.psi, f.slip_shift,
                     dev_FV(c), dev_Fd(c), f.Ys[s], f.Yp[s], c->d[0].u,
                     c->cfg.nk, c->lv[0].K, c->lv[0].P, go);
}

void launch_fault_update(qdeq_ctx *c, int s) {
  FaultBufs &f = c->fb;
  hipLaunchKernelGGL(k_fault_update, nodes_grid(c->cfg.nk), dim3(FT), 0,
                     c->stream, c->d[0].u, c->lv[0].P, c->cfg.nk, f.mu_f,
                     f.tau_n0, f.Yp[s], f.a, f.b, f.eta, c->cfg.V0, c->cfg.f0,
                     c->cfg.L, c->c_stress[0], c->c_stress[1], c->c_stress[2],
                     c->c_stress[3], f.Ts, f.Tn, f.FV[s], f.Fd[s], c->flags);
}

void launch_update_stage(qdeq_ctx *c, int s, const int *spec) {
  FaultBufs &f = c->fb;
  const UpdateArgs ua{f.mu_f, f.tau_n0, f.Yp[s], f.a, f.b, f.eta, c->cfg.V0,
                      c->cfg.f0, c->cfg.L, c->c_stress[0], c->c_stress[1],
                      c->c_stress[2], c->c_stress[3], f.Ts, f.Tn, f.FV[s], f.Fd[s],
                      c->flags};
  const StageArgs sa{s + 1, c->scal + SC_T, kC[s + 1], c->cfg.Vini, f.slip, f.psi,
                     f.slip_shift, dev_FV(c), dev_Fd(c), f.Ys[s + 1], f.Yp[s + 1]};
  hipLaunchKernelGGL(k_update_stage, nodes_grid(c->cfg.nk), dim3(FT), 0, c->stream,
                     c->d[0].u, c->cfg.nk, c->lv[0].K, c->lv[0].P, ua, sa, spec);
}

void launch_update_next(qdeq_ctx *c, double t, double dt, const int *spec) {
  FaultBufs &f = c->fb;
  const UpdateArgs ua{f.mu_f, f.tau_n0, f.Yp[1], f.a, f.b, f.eta, c->cfg.V0,
                      c->cfg.f0, c->cfg.L, c->c_stress[0], c->c_stress[1],
                      c->c_stress[2], c->c_stress[3], f.Ts, f.Tn, f.FV[1], f.Fd[1],
                      c->flags};
  const StageArgs sa{2, c->scal + SC_T, kC[2], c->cfg.Vini, f.slip, f.psi,
                     f.slip_shift, dev_FV(c), dev_Fd(c), f.Ys[2], f.Yp[2]};
  hipLaunchKernelGGL(k_update_next, nodes_grid(c->nz), dim3(FT), 0, c->stream,
                     c->d[0].u, c->cfg.nk, c->nz, c->lv[0].K, c->lv[0].P, ua, sa, t,
                     dt, c->scal + SC_T, spec);
}

void launch_zero_stage_rates(qdeq_ctx *c, double t, double dt) {
  hipLaunchKernelGGL(k_zero_rates, nodes_grid(c->nz), dim3(FT), 0, c->stream,
                     dev_FV(c), dev_Fd(c), c->nz, c->scal + SC_T, t, dt);
}

void launch_combine(qdeq_ctx *c) {
  FaultBufs &f = c->fb;
  hipLaunchKernelGGL(k_combine, dim3(1), dim3(1024), 0, c->stream, c->scal + SC_T, f.slip,
                     f.psi, dev_FV(c), dev_Fd(c), f.y3s, f.y3p, f.y4s, f.y4p,
                     c->cfg.nk, c->scal + SC_TOTERR, c->gate, c->cfg.totTol);
}

void launch_commit(qdeq_ctx *c, double t_new, const int *go) {
  FaultBufs &f = c->fb;
  hipLaunchKernelGGL(k_commit, nodes_grid(c->cfg.nk), dim3(FT), 0, c->stream,
                     f.slip, f.psi, f.y4s, f.y4p, f.FV[1], f.Fd[1], c->cfg.nk,
                     c->scal + SC_T, t_new, go);
}

void launch_maxima(qdeq_ctx *c) {
  hipLaunchKernelGGL(k_maxima, dim3(1), dim3(1024), 0, c->stream, c->fb.FV[1],
                     c->fb.Ts, c->cfg.nk, c->scal + SC_VMAX);
}

}  // namespace qdeq
