/*
 * device_path.c -- the drop-in entry points of include/qdeq.h called from C
 * the way a reference maintainer would (Option A): once with
 * host arrays (QDEQ_HOST) and once with device buffers (QDEQ_DEVICE) that
 * the caller filled with asynchronous copies on the null stream right
 * before each call.  The two sets of results must be identical bit for bit.
 * The device inputs start out as NaN, so a call that read them before the
 * caller's copies landed would differ.
 *
 * The problem: 65 x 65 nodes, 750 m spacing, uniform mu = 32 GPa, 3
 * levels, the fast cycle to a 1e-10 relative residual; the fault and remote
 * columns hold a smooth slip profile and a uniform far-field offset.
 *
 * usage: device_path [device]; prints "device path OK" and exits 0.
 */
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "qdeq.h"

#define HIPCK(call)                                                        \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess) {                                                \
      fprintf(stderr, "%s: %s\n", #call, hipGetErrorString(e_));           \
      exit(2);                                                             \
    }                                                                      \
  } while (0)

#define QCK(q, call)                                                       \
  do {                                                                     \
    int s_ = (call);                                                       \
    if (s_ != QDEQ_OK) {                                                   \
      fprintf(stderr, "%s: status %d: %s\n", #call, s_,                    \
              (q) ? qdeq_last_error(q) : qdeq_global_error());             \
      exit(3);                                                             \
    }                                                                      \
  } while (0)

static double *dev_nan(size_t n) {
  double *p = NULL;
  HIPCK(hipMalloc((void **)&p, n * sizeof(double)));
  HIPCK(hipMemset(p, 0xff, n * sizeof(double))); /* NaN everywhere */
  return p;
}

/* staging copy on the device, then an asynchronous copy into dst on the
 * null stream that the next library call must wait for */
static void put_async(double *dst, double *stage, const double *src, size_t n) {
  HIPCK(hipMemcpy(stage, src, n * sizeof(double), hipMemcpyHostToDevice));
  HIPCK(hipMemcpyAsync(dst, stage, n * sizeof(double), hipMemcpyDeviceToDevice, 0));
}

static int same(const char *what, const double *a, const double *b, size_t n) {
  if (memcmp(a, b, n * sizeof(double)) != 0) {
    for (size_t i = 0; i < n; i++)
      if (memcmp(a + i, b + i, sizeof(double)) != 0) {
        fprintf(stderr, "%s differs at %zu: %.17g vs %.17g\n", what, i, a[i], b[i]);
        break;
      }
    return 0;
  }
  return 1;
}

int main(int argc, char **argv) {
  const int device = argc > 1 ? atoi(argv[1]) : 0;
  const int ni = 65, nk = 65, nx = ni + 2, nz = nk + 2;
  const size_t ng = (size_t)nx * nz;
  qdeq_config cfg;
  qdeq_default_config(&cfg);
  cfg.ni = ni;
  cfg.nk = nk;
  cfg.dh = 750.0;
  cfg.n_levels = 3;
  cfg.cycle = QDEQ_CYCLE_FAST;
  cfg.n_pre = 1;
  cfg.n_post = 3;
  cfg.n_coarse = 0;
  cfg.relax = 1.15;
  cfg.tol_mode = QDEQ_TOL_REL;
  cfg.tol = 1e-10;
  cfg.device = device;

  double *mu = calloc(ng, sizeof(double)), *U0 = calloc(ng, sizeof(double));
  double *slip = calloc(nz, sizeof(double)), *ss = calloc(nz, sizeof(double));
  double *tau_n0 = calloc(nz, sizeof(double)), *psi = calloc(nz, sizeof(double));
  double *a = calloc(nz, sizeof(double)), *b = calloc(nz, sizeof(double));
  double *eta = calloc(nz, sizeof(double));
  for (int k = 1; k <= nk; k++)
    for (int i = 1; i <= ni; i++) {
      mu[(size_t)k * nx + i] = 32.0;
      U0[(size_t)k * nx + i] = 39.8 * (i - 1) / (double)(ni - 1) +
                               0.3 * sin(0.11 * i) * cos(0.07 * k);
    }
  for (int k = 1; k <= nk; k++) {
    const double d = (nk - k) * cfg.dh * 1e-3; /* depth, km */
    slip[k] = 2.0 * exp(-((d - 8.0) * (d - 8.0)) / 20.0);
    ss[k] = 39.8;
    tau_n0[k] = -50.0;
    psi[k] = 0.6 + 0.015 * log(1e-6 / 1e-9) + 0.01 * sin(0.2 * k);
    a[k] = d <= 15.0 ? 0.010 : 0.025;
    b[k] = 0.015;
    eta[k] = 32.0 / (2.0 * 3.464);
  }
  ss[0] = ss[nz - 1] = 39.8;

  qdeq_ctx *q = NULL;
  QCK(NULL, qdeq_create(&cfg, mu, NULL, &q));
  const double t = 3.0e8, ratio = 0.62, dt = 2.5e6;

  /* host path (arrays as the reference keeps them) */
  double *Uh = malloc(ng * sizeof(double));
  memcpy(Uh, U0, ng * sizeof(double));
  double *Tsh = calloc(nz, sizeof(double)), *Tnh = calloc(nz, sizeof(double));
  double *Vh = calloc(nz, sizeof(double)), *dph = calloc(nz, sizeof(double));
  int vch = 0;
  double r0h = 0, rh = 0;
  QCK(q, qdeq_set_boundary(q, Uh, slip, ss, t, ratio, dt, QDEQ_HOST));
  QCK(q, qdeq_ode_solve(q, Uh, QDEQ_HOST, &vch, &r0h, &rh));
  QCK(q, qdeq_eval_stress(q, Uh, mu, tau_n0, Tsh, Tnh, QDEQ_HOST));
  QCK(q, qdeq_nr_search(q, Tsh, Tnh, psi, a, eta, Vh, QDEQ_HOST));
  QCK(q, qdeq_cal_dpsi(q, psi, Vh, b, dph, QDEQ_HOST));

  /* device path: the same calls on device buffers */
  double *dU = dev_nan(ng), *dmu = dev_nan(ng), *stg = dev_nan(ng);
  double *dslip = dev_nan(nz), *dss = dev_nan(nz), *dtn = dev_nan(nz);
  double *dpsi = dev_nan(nz), *da = dev_nan(nz), *db = dev_nan(nz);
  double *deta = dev_nan(nz), *dTs = dev_nan(nz), *dTn = dev_nan(nz);
  double *dV = dev_nan(nz), *ddp = dev_nan(nz);
  double *s1 = dev_nan(nz), *s2 = dev_nan(nz);
  put_async(dU, stg, U0, ng);
  put_async(dslip, s1, slip, nz);
  put_async(dss, s2, ss, nz);
  QCK(q, qdeq_set_boundary(q, dU, dslip, dss, t, ratio, dt, QDEQ_DEVICE));
  int vcd = 0;
  double r0d = 0, rd = 0;
  QCK(q, qdeq_ode_solve(q, dU, QDEQ_DEVICE, &vcd, &r0d, &rd));
  double *s3 = dev_nan(nz);
  put_async(dmu, stg, mu, ng);
  put_async(dtn, s3, tau_n0, nz);
  QCK(q, qdeq_eval_stress(q, dU, dmu, dtn, dTs, dTn, QDEQ_DEVICE));
  double *s4 = dev_nan(nz), *s5 = dev_nan(nz), *s6 = dev_nan(nz), *s7 = dev_nan(nz);
  put_async(dpsi, s4, psi, nz);
  put_async(da, s5, a, nz);
  put_async(deta, s6, eta, nz);
  HIPCK(hipMemsetAsync(dV, 0, nz * sizeof(double), 0));
  QCK(q, qdeq_nr_search(q, dTs, dTn, dpsi, da, deta, dV, QDEQ_DEVICE));
  put_async(db, s7, b, nz);
  QCK(q, qdeq_cal_dpsi(q, dpsi, dV, db, ddp, QDEQ_DEVICE));

  double *Ud = malloc(ng * sizeof(double));
  double *Tsd = malloc(nz * sizeof(double)), *Tnd = malloc(nz * sizeof(double));
  double *Vd = malloc(nz * sizeof(double)), *dpd = malloc(nz * sizeof(double));
  HIPCK(hipMemcpy(Ud, dU, ng * sizeof(double), hipMemcpyDeviceToHost));
  HIPCK(hipMemcpy(Tsd, dTs, nz * sizeof(double), hipMemcpyDeviceToHost));
  HIPCK(hipMemcpy(Tnd, dTn, nz * sizeof(double), hipMemcpyDeviceToHost));
  HIPCK(hipMemcpy(Vd, dV, nz * sizeof(double), hipMemcpyDeviceToHost));
  HIPCK(hipMemcpy(dpd, ddp, nz * sizeof(double), hipMemcpyDeviceToHost));

  int ok = 1;
  ok &= same("U", Uh, Ud, ng);
  ok &= same("Ts", Tsh + 1, Tsd + 1, nk);
  ok &= same("Tn", Tnh + 1, Tnd + 1, nk);
  ok &= same("V", Vh + 1, Vd + 1, nk);
  ok &= same("dpsi", dph + 1, dpd + 1, nk);
  if (vch != vcd || r0h != r0d || rh != rd) {
    fprintf(stderr, "solve outputs differ: %d/%d %.17g/%.17g %.17g/%.17g\n", vch, vcd,
            r0h, r0d, rh, rd);
    ok = 0;
  }
  if (!(vch >= 1 && rh <= 1e-10 * r0h)) {
    fprintf(stderr, "solve did not converge: %d V-cycles, %g / %g\n", vch, rh, r0h);
    ok = 0;
  }
  /* an invalid pointer space is refused before anything runs */
  if (qdeq_ode_solve(q, Uh, 7, NULL, NULL, NULL) != QDEQ_ERR_ARG) {
    fprintf(stderr, "pointer space 7 accepted\n");
    ok = 0;
  }
  qdeq_destroy(q);
  if (!ok) return 1;
  printf("device path OK: %d V-cycles, residual %.3e of %.3e\n", vcd, rd, r0d);
  return 0;
}
