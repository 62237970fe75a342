/*
 * qdeq_oracle.c -- CPU FP64 restatement of the reference solver.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): see
 * qdeq_oracle.h.  Built with gcc -O2 -ffp-contract=off so every expression is
 * evaluated in the order written, as the reference writes it.
 *
 * Reference files cited below are from XuSun-SUSTech/QDEqCycle_FDM.
 */
#include "qdeq_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OMIN(a, b) (((a) <= (b)) ? (a) : (b)) /* header.h:31 */
#define OMAX(a, b) (((a) >= (b)) ? (a) : (b)) /* header.h:30 */

/* ------------------------------------------------------------------------ */
/* Equivalent medium: odeSolver.cpp:134-268 (Lagrange interpolation) and     */
/* 337-589 (harmonic mean of 400 interpolated samples per face).            */
/* ------------------------------------------------------------------------ */

/* interp_2d / interp2_2d (odeSolver.cpp:134-200, 202-268): tensor Lagrange
 * interpolation of z[ni][nj] (row-major, z[i*nj+j]) at (xi, yi).  Products
 * and sums are taken in the reference's order. */
static double orc_lagrange_2d(const double *x, const double *y,
                              const double *z, int ni, int nj, double xi,
                              double yi) {
  double Lx[4], xt[4], xb[4], Ly[4], yt[4], yb[4];
  int i, j, k;
  for (k = 0; k < ni; k++) Lx[k] = 1.0;
  for (i = 0; i < ni; i++) {
    for (k = 0; k < ni; k++) xb[k] = x[i] - x[k];
    xb[i] = 1.0;
    for (k = 0; k < ni; k++) xt[k] = xi - x[k];
    xt[i] = 1.0;
    for (j = 0; j < ni; j++) Lx[i] = Lx[i] * xt[j] / xb[j];
  }
  for (k = 0; k < nj; k++) Ly[k] = 1.0;
  for (i = 0; i < nj; i++) {
    for (k = 0; k < nj; k++) yb[k] = y[i] - y[k];
    yb[i] = 1.0;
    for (k = 0; k < nj; k++) yt[k] = yi - y[k];
    yt[i] = 1.0;
    for (j = 0; j < nj; j++) Ly[i] = Ly[i] * yt[j] / yb[j];
  }
  double zi = 0.0;
  for (j = 0; j < nj; j++)
    for (i = 0; i < ni; i++) zi = zi + Lx[i] * Ly[j] * z[i * nj + j];
  return zi;
}

/* 20x20 sample harmonic mean (odeSolver.cpp:417-419, 442-460). */
static double orc_face_harmonic(const double *xn, const double *yn,
                                const double *m, int ni, int nj) {
  const int NGRIDX = 10, NGRIDY = 10;
  const int nsamp = (2 * NGRIDX) * (2 * NGRIDY);
  double mu0 = 0.0;
  for (int mj = -1 * NGRIDY; mj <= NGRIDY; mj++) {
    if (mj == 0) continue;
    for (int mi = -1 * NGRIDX; mi <= NGRIDX; mi++) {
      if (mi == 0) continue;
      double MI = mi / ((NGRIDX + 1.0) * 2.0);
      double MJ = mj / ((NGRIDY + 1.0) * 2.0);
      double mu_interp = orc_lagrange_2d(xn, yn, m, ni, nj, MI, MJ);
      mu0 = mu0 + 1.0 / mu_interp;
    }
  }
  return nsamp / mu0;
}

/* SetMedium, odeSolver.cpp:337-589.  mu is the nodal nx*nz array; mu_equi
 * is the doubled (2nx-1)*(2nz-1) array.  Entries the reference never writes
 * stay 0 (pMedium is memset, odeSolver.cpp:75).
 *
 * Quirk: the x-face loop never assigns mu_matrix[1][1] (odeSolver.cpp:434-436
 * skip it), so the reference reads an uninitialised stack value there.  The
 * restatement uses the evident intent, the nodal value mu_half[index3 + 2]
 * (row k, node i-1), which is what every other entry of the 3x4 block
 * follows.  The mu_half half-points (odeSolver.cpp:381-404) are never read by
 * the face loops and are not formed. */
void orc_mu_equi(int ni, int nk, const double *mu, double *me) {
  const int nx = ni + 2, nz = nk + 2;
  const long W2 = 2L * nx - 1, H2 = 2L * nz - 1;
  const int ni1 = 2, ni2 = ni + 1, nk1 = 2, nk2 = nk + 1; /* 1-based bounds */
  double *mh = (double *)calloc((size_t)(W2 * H2), sizeof(double));
  long i, k, index, index2, index3, index4, index5;
  memset(me, 0, sizeof(double) * (size_t)(W2 * H2));

  for (k = nk1 - 1; k < nk2; k++) /* 337-346 */
    for (i = ni1 - 1; i < ni2; i++) {
      index = i + k * nx;
      index2 = i * 2 + (k * 2) * W2;
      mh[index2] = mu[index];
      me[index2] = mu[index];
    }
  for (i = ni1 - 1; i < ni2; i++) { /* 347-353 */
    index = i * 2 + 0 * W2;
    index2 = i * 2 + ((nk1 - 1) * 2) * W2;
    mh[index] = mh[index2];
    me[index] = mh[index2];
  }
  for (i = ni1 - 1; i < ni2; i++) { /* 354-360 */
    index = i * 2 + ((nz - 1) * 2) * W2;
    index2 = i * 2 + ((nk2 - 1) * 2) * W2;
    mh[index] = mh[index2];
    me[index] = mh[index2];
  }
  for (k = 0; k < nz; k++) { /* 362-368 */
    index = 0 * 2 + (k * 2) * W2;
    index2 = (ni1 - 1) * 2 + (k * 2) * W2;
    mh[index] = mh[index2];
    me[index] = mh[index2];
  }
  for (k = 0; k < nz; k++) { /* 370-376 */
    index = (nx - 1) * 2 + (k * 2) * W2;
    index2 = (ni2 - 1) * 2 + (k * 2) * W2;
    mh[index] = mh[index2];
    me[index] = mh[index2];
  }

  static const double x_index[3] = {-1.0, 0.0, 1.0};          /* 410 */
  static const double y_index[4] = {-1.5, -0.5, 0.5, 1.5};    /* 411 */
  static const double x_index2[4] = {-1.5, -0.5, 0.5, 1.5};   /* 414 */
  static const double y_index2[3] = {-1.0, 0.0, 1.0};         /* 415 */
  double m[3 * 4], m2[4 * 3];

  /* x-faces between nodes i-1 and i (odeSolver.cpp:421-462) */
  for (k = nk1 - 1; k < nk2; k++)
    for (i = ni1; i < ni2; i++) {
      index = (i * 2 - 1) + (k * 2) * W2;
      index2 = (i - 2) * 2 + ((k - 1) * 2) * W2;
      index3 = (i - 2) * 2 + (k * 2) * W2;
      index4 = (i - 2) * 2 + ((k + 1) * 2) * W2;
      m[0 * 4 + 0] = mh[index2];
      m[0 * 4 + 1] = mh[index2 + 2];
      m[0 * 4 + 2] = mh[index2 + 4];
      m[0 * 4 + 3] = mh[index2 + 6];
      m[1 * 4 + 0] = mh[index3];
      m[1 * 4 + 1] = mh[index3 + 2]; /* unassigned in the reference */
      m[1 * 4 + 2] = mh[index3 + 4];
      m[1 * 4 + 3] = mh[index3 + 6];
      m[2 * 4 + 0] = mh[index4];
      m[2 * 4 + 1] = mh[index4 + 2];
      m[2 * 4 + 2] = mh[index4 + 4];
      m[2 * 4 + 3] = mh[index4 + 6];
      me[index] = orc_face_harmonic(x_index, y_index, m, 3, 4);
    }

  /* fills the 4x3 block for a z-face from four nodal rows (464-506) */
#define ORC_FILL_M2(r0, r1, r2, r3)            \
  do {                                         \
    m2[0 * 3 + 0] = mh[(r0)];                  \
    m2[0 * 3 + 1] = mh[(r0) + 2];              \
    m2[0 * 3 + 2] = mh[(r0) + 4];              \
    m2[1 * 3 + 0] = mh[(r1)];                  \
    m2[1 * 3 + 1] = mh[(r1) + 2];              \
    m2[1 * 3 + 2] = mh[(r1) + 4];              \
    m2[2 * 3 + 0] = mh[(r2)];                  \
    m2[2 * 3 + 1] = mh[(r2) + 2];              \
    m2[2 * 3 + 2] = mh[(r2) + 4];              \
    m2[3 * 3 + 0] = mh[(r3)];                  \
    m2[3 * 3 + 1] = mh[(r3) + 2];              \
    m2[3 * 3 + 2] = mh[(r3) + 4];              \
  } while (0)

  /* z-faces between rows k-1 and k (odeSolver.cpp:464-506) */
  for (k = nk1; k < nk2; k++)
    for (i = ni1; i < ni2 - 1; i++) {
      index = i * 2 + (k * 2 - 1) * W2;
      index2 = (i - 1) * 2 + ((k - 2) * 2) * W2;
      index3 = (i - 1) * 2 + ((k - 1) * 2) * W2;
      index4 = (i - 1) * 2 + (k * 2) * W2;
      index5 = (i - 1) * 2 + ((k + 1) * 2) * W2;
      ORC_FILL_M2(index2, index3, index4, index5);
      me[index] = orc_face_harmonic(x_index2, y_index2, m2, 4, 3);
    }
  /* bottom boundary z-face (odeSolver.cpp:509-548) */
  for (i = ni1; i < ni2 - 1; i++) {
    index = i * 2 + ((nk1 - 1) * 2 - 1) * W2;
    index2 = (i - 1) * 2 + 0 * W2;
    index3 = (i - 1) * 2 + 0 * W2;
    index4 = (i - 1) * 2 + ((nk1 - 1) * 2) * W2;
    index5 = (i - 1) * 2 + (nk1 * 2) * W2;
    ORC_FILL_M2(index2, index3, index4, index5);
    me[index] = orc_face_harmonic(x_index2, y_index2, m2, 4, 3);
  }
  /* top boundary z-face (odeSolver.cpp:550-589) */
  for (i = ni1; i < ni2 - 1; i++) {
    index = i * 2 + (nk2 * 2 - 1) * W2;
    index2 = (i - 1) * 2 + ((nk2 - 2) * 2) * W2;
    index3 = (i - 1) * 2 + ((nk2 - 1) * 2) * W2;
    index4 = (i - 1) * 2 + (nk2 * 2) * W2;
    index5 = (i - 1) * 2 + (nk2 * 2) * W2;
    ORC_FILL_M2(index2, index3, index4, index5);
    me[index] = orc_face_harmonic(x_index2, y_index2, m2, 4, 3);
  }
#undef ORC_FILL_M2
  free(mh);
}

/* ------------------------------------------------------------------------ */
/* Multigrid (odeSolver.cpp:652-1147)                                        */
/* ------------------------------------------------------------------------ */

/* Threads: the loops over the points of one colour, of a residual, a
 * restriction or a prolongation update independent points, so they are
 * split over OpenMP threads by rows without changing any result; sums (the
 * residual norm) stay sequential.  Grids below ORC_OMP_MIN points run on
 * one thread. */
#define ORC_OMP_MIN 16384
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

/* Level hierarchy, odeSolver.cpp:1384-1405: N=nk, K=ni at level 0,
 * N <- (N-1)/2+1, dh doubled, fin_coa = 2^level. */
void orc_level_dims(int ni, int nk, double dh, int n_levels, int *N, int *K,
                    int *fin_coa, double *DH) {
  int n = nk, k = ni, fc = 1;
  double h = dh;
  N[0] = n; K[0] = k; fin_coa[0] = 1; DH[0] = h;
  for (int l = 0; l < n_levels - 1; l++) {
    n = (n - 1) / 2 + 1;
    k = (k - 1) / 2 + 1;
    h *= 2;
    fc *= 2;
    N[l + 1] = n; K[l + 1] = k; fin_coa[l + 1] = fc; DH[l + 1] = h;
  }
}

/* Face-coefficient indices into mu_equi for compact point (j,k) at a level:
 * level 0 reads the equivalent-medium faces (odeSolver.cpp:785-789), coarse
 * levels read nodal mu at the fine node under each coarse-face midpoint
 * (odeSolver.cpp:823-827). */
static void orc_mu_index(int level, int fc, long W2, long j, long k, long *i1,
                         long *i2, long *i3, long *i4) {
  if (level == 0) {
    *i1 = ((k + 1) * 2 - 1) + ((j + 1) * 2) * W2;
    *i2 = ((k + 1) * 2 + 1) + ((j + 1) * 2) * W2;
    *i3 = (k + 1) * 2 + ((j + 1) * 2 - 1) * W2;
    *i4 = (k + 1) * 2 + ((j + 1) * 2 + 1) * W2;
  } else {
    *i1 = ((k * 2 - 1) * fc / 2 + 1) * 2 + ((j * fc + 1) * 2) * W2;
    *i2 = ((k * 2 + 1) * fc / 2 + 1) * 2 + ((j * fc + 1) * 2) * W2;
    *i3 = (k * fc + 1) * 2 + (((j * 2 - 1) * fc / 2 + 1) * 2) * W2;
    *i4 = (k * fc + 1) * 2 + (((j * 2 + 1) * fc / 2 + 1) * 2) * W2;
  }
}

/* rbgs_kernel<phase>, odeSolver.cpp:757-861.  A point is updated when
 * (j+k)%2 == phase; the update reads only the other colour, so the loop
 * order is immaterial.  Top/bottom rows mirror the vertical neighbour; on
 * coarse levels they also reuse the single available vertical coefficient
 * twice (830-847). */
void orc_rbgs(double *u, const double *f, const double *mu, int ni, int nk,
              double dh, int level, int phase, double relax) {
  int N[ORC_MAX_LEVELS], K[ORC_MAX_LEVELS], FC[ORC_MAX_LEVELS];
  double DH[ORC_MAX_LEVELS];
  orc_level_dims(ni, nk, dh, level + 1, N, K, FC, DH);
  const long Nn = N[level], Kk = K[level], W2 = 2L * (ni + 2) - 1;
  const double DH2 = DH[level] * DH[level];
#pragma omp parallel for schedule(static) if (Nn * Kk > ORC_OMP_MIN)
  for (long j = 0; j < Nn; j++)
    for (long k = 1; k < Kk - 1; k++) {
      if ((0 + j + k) % 2 != phase) continue;
      long i1, i2, i3, i4;
      orc_mu_index(level, FC[level], W2, j, k, &i1, &i2, &i3, &i4);
      const long c = k + j * Kk;
      const double uW = u[c - 1], uE = u[c + 1];
      double m1 = mu[i1], m2 = mu[i2], m3, m4, uS, uN;
      if (j == 0) {
        uS = u[c + Kk];
        uN = u[c + Kk];
        m4 = mu[i4];
        m3 = (level == 0) ? mu[i3] : m4;
      } else if (j == Nn - 1) {
        uS = u[c - Kk];
        uN = u[c - Kk];
        m3 = mu[i3];
        m4 = (level == 0) ? mu[i4] : m3;
      } else {
        uS = u[c - Kk];
        uN = u[c + Kk];
        m3 = mu[i3];
        m4 = mu[i4];
      }
      u[c] = (relax / (m1 + m2 + m3 + m4)) *
                 (m1 * uW + m2 * uE + m3 * uS + m4 * uN - DH2 * f[c]) +
             (1.00 - relax) * u[c];
    }
}

/* residual_kernel, odeSolver.cpp:863-951 (and calcuDelta_dev 1045-1099 at
 * level 0).  Entries at k=0 and k=K-1 are not written. */
void orc_residual(double *r, const double *u, const double *f,
                  const double *mu, int ni, int nk, double dh, int level) {
  int N[ORC_MAX_LEVELS], K[ORC_MAX_LEVELS], FC[ORC_MAX_LEVELS];
  double DH[ORC_MAX_LEVELS];
  orc_level_dims(ni, nk, dh, level + 1, N, K, FC, DH);
  const long Nn = N[level], Kk = K[level], W2 = 2L * (ni + 2) - 1;
  const double DH2 = DH[level] * DH[level];
#pragma omp parallel for schedule(static) if (Nn * Kk > ORC_OMP_MIN)
  for (long j = 0; j < Nn; j++)
    for (long k = 1; k < Kk - 1; k++) {
      long i1, i2, i3, i4;
      orc_mu_index(level, FC[level], W2, j, k, &i1, &i2, &i3, &i4);
      const long c = k + j * Kk;
      double m1 = mu[i1], m2 = mu[i2], m3, m4, uS, uN;
      if (j == 0) {
        uS = u[c + Kk];
        uN = u[c + Kk];
        m4 = mu[i4];
        m3 = (level == 0) ? mu[i3] : m4;
      } else if (j == Nn - 1) {
        uS = u[c - Kk];
        uN = u[c - Kk];
        m3 = mu[i3];
        m4 = (level == 0) ? mu[i4] : m3;
      } else {
        uS = u[c - Kk];
        uN = u[c + Kk];
        m3 = mu[i3];
        m4 = mu[i4];
      }
      r[c] = f[c] - (m1 * u[c - 1] + m2 * u[c + 1] + m3 * uS + m4 * uN -
                     (m1 + m2 + m3 + m4) * u[c]) /
                        DH2;
    }
}

/* restriction_kernel, odeSolver.cpp:953-1015 (level = coarse level).
 * Full weighting with the first/last coarse rows mirroring the fine row
 * inside the domain; also zeroes the coarse correction e2. */
void orc_restrict(double *r2, const double *r, double *e2, int ni, int nk,
                  int level) {
  int N[ORC_MAX_LEVELS], K[ORC_MAX_LEVELS], FC[ORC_MAX_LEVELS];
  double DH[ORC_MAX_LEVELS];
  orc_level_dims(ni, nk, 1.0, level + 1, N, K, FC, DH);
  const long Nc = N[level], Kc = K[level], Kf = K[level - 1];
#define RF(jj, kk) r[(long)(kk) + (long)(jj) * Kf]
#pragma omp parallel for schedule(static) if (Nc * Kc > ORC_OMP_MIN)
  for (long j = 0; j < Nc; j++)
    for (long k = 1; k < Kc - 1; k++) {
      const long c = k + j * Kc;
      if (j == 0) {
        r2[c] = 0.06250 * (4.0 * RF(2 * j, 2 * k) + 2.0 * RF(2 * j, 2 * k - 1) +
                           2.0 * RF(2 * j, 2 * k + 1) +
                           2.0 * RF(2 * j + 1, 2 * k) +
                           2.0 * RF(2 * j + 1, 2 * k) + RF(2 * j + 1, 2 * k - 1) +
                           RF(2 * j + 1, 2 * k + 1) + RF(2 * j + 1, 2 * k - 1) +
                           RF(2 * j + 1, 2 * k + 1));
      } else if (j == Nc - 1) {
        r2[c] = 0.06250 * (4.0 * RF(2 * j, 2 * k) + 2.0 * RF(2 * j, 2 * k - 1) +
                           2.0 * RF(2 * j, 2 * k + 1) +
                           2.0 * RF(2 * j - 1, 2 * k) +
                           2.0 * RF(2 * j - 1, 2 * k) + RF(2 * j - 1, 2 * k - 1) +
                           RF(2 * j - 1, 2 * k + 1) + RF(2 * j - 1, 2 * k - 1) +
                           RF(2 * j - 1, 2 * k + 1));
      } else {
        r2[c] = 0.06250 * (4.0 * RF(2 * j, 2 * k) + 2.0 * RF(2 * j, 2 * k - 1) +
                           2.0 * RF(2 * j, 2 * k + 1) +
                           2.0 * RF(2 * j - 1, 2 * k) +
                           2.0 * RF(2 * j + 1, 2 * k) + RF(2 * j - 1, 2 * k - 1) +
                           RF(2 * j - 1, 2 * k + 1) + RF(2 * j + 1, 2 * k - 1) +
                           RF(2 * j + 1, 2 * k + 1));
      }
      e2[c] = 0.0;
    }
#undef RF
}

/* prolongate_kernel, odeSolver.cpp:1017-1043 (level = coarse level).
 * Bilinear correction over coarse j < Nc-1, k < Kc-1: the last fine row and
 * column receive no correction.  With full_top (the engine's fast cycle,
 * no reference counterpart) the last coarse row also corrects the last fine
 * row. */
void orc_prolong_full(double *u, const double *e2, int ni, int nk, int level,
                      int full_top) {
  int N[ORC_MAX_LEVELS], K[ORC_MAX_LEVELS], FC[ORC_MAX_LEVELS];
  double DH[ORC_MAX_LEVELS];
  orc_level_dims(ni, nk, 1.0, level + 1, N, K, FC, DH);
  const long Nc = N[level], Kc = K[level], Kf = K[level - 1];
#define UF(jj, kk) u[(long)(kk) + (long)(jj) * Kf]
#define EC(jj, kk) e2[(long)(kk) + (long)(jj) * Kc]
#pragma omp parallel for schedule(static) if (Nc * Kc > ORC_OMP_MIN)
  for (long j = 0; j < Nc - 1; j++)
    for (long k = 0; k < Kc - 1; k++) {
      UF(2 * j, 2 * k) += EC(j, k);
      UF(2 * j, 2 * k + 1) += 0.5 * (EC(j, k) + EC(j, k + 1));
      UF(2 * j + 1, 2 * k) += 0.5 * (EC(j, k) + EC(j + 1, k));
      UF(2 * j + 1, 2 * k + 1) +=
          0.25 * (EC(j, k) + EC(j, k + 1) + EC(j + 1, k) + EC(j + 1, k + 1));
    }
  if (full_top) {
    const long j = Nc - 1;
    for (long k = 0; k < Kc - 1; k++) {
      UF(2 * j, 2 * k) += EC(j, k);
      UF(2 * j, 2 * k + 1) += 0.5 * (EC(j, k) + EC(j, k + 1));
    }
  }
#undef UF
#undef EC
}

void orc_prolong(double *u, const double *e2, int ni, int nk, int level) {
  orc_prolong_full(u, e2, ni, nk, level, 0);
}

/* Euclidean norm standing in for cublasDnrm2 (odeSolver.cpp:694,700); the
 * reference's library routine has no published summation order, so only
 * the stop test depends on it.  Pinned by closed forms, exact scaling by
 * powers of two and the rounding bound of a sequential sum of n positive
 * terms (tests/test_oracle.py). */
double orc_nrm2(const double *x, long n) {
  double s = 0.0;
  for (long i = 0; i < n; i++) s += x[i] * x[i];
  return sqrt(s);
}

struct orc_mg {
  int ni, nk, L, nu;
  int n_pre, n_post, n_coarse, full_top;  /* cycle shape (reference: nu,nu,nu,0) */
  double *cinv;                  /* fast cycle, n_coarse == 0: inverse of the */
  int cn;                        /* coarsest operator (row-major), unknowns   */
  double dh, relax;
  int N[ORC_MAX_LEVELS], K[ORC_MAX_LEVELS], FC[ORC_MAX_LEVELS];
  double DH[ORC_MAX_LEVELS];
  double *mu;                     /* mu_equi copy */
  double *res[ORC_MAX_LEVELS];    /* level l residual, n_l       (1393) */
  double *res2[ORC_MAX_LEVELS];   /* restricted residual, n_{l+1} (1399) */
  double *err2[ORC_MAX_LEVELS];   /* coarse correction, n_{l+1}  (1400) */
  double *u, *f, *delta, *tmp;    /* level-0 compact u, f (zero), delta_res */
};

orc_mg *orc_mg_create(int ni, int nk, double dh, int n_levels, int n_cycles,
                      double relax, const double *mu_equi) {
  if (n_levels < 1 || n_levels > ORC_MAX_LEVELS) return NULL;
  orc_mg *mg = (orc_mg *)calloc(1, sizeof(orc_mg));
  mg->ni = ni; mg->nk = nk; mg->L = n_levels; mg->nu = n_cycles;
  mg->n_pre = mg->n_post = mg->n_coarse = n_cycles;
  mg->full_top = 0;
  mg->dh = dh; mg->relax = relax;
  orc_level_dims(ni, nk, dh, n_levels, mg->N, mg->K, mg->FC, mg->DH);
  const size_t nmu = (size_t)(2 * (ni + 2) - 1) * (size_t)(2 * (nk + 2) - 1);
  mg->mu = (double *)malloc(nmu * sizeof(double));
  memcpy(mg->mu, mu_equi, nmu * sizeof(double));
  for (int l = 0; l < n_levels - 1; l++) {
    mg->res[l] = (double *)calloc((size_t)mg->N[l] * mg->K[l], sizeof(double));
    size_t nc = (size_t)mg->N[l + 1] * mg->K[l + 1];
    mg->res2[l] = (double *)calloc(nc, sizeof(double));
    mg->err2[l] = (double *)calloc(nc, sizeof(double));
  }
  size_t n0 = (size_t)ni * nk;
  mg->u = (double *)calloc(n0, sizeof(double));
  mg->f = (double *)calloc(n0, sizeof(double)); /* never written (Q1) */
  mg->delta = (double *)calloc(n0, sizeof(double));
  mg->tmp = (double *)calloc(n0, sizeof(double));
  return mg;
}

void orc_mg_destroy(orc_mg *mg) {
  if (!mg) return;
  for (int l = 0; l < mg->L - 1; l++) {
    free(mg->res[l]); free(mg->res2[l]); free(mg->err2[l]);
  }
  free(mg->mu); free(mg->u); free(mg->f); free(mg->delta); free(mg->tmp);
  free(mg->cinv);
  free(mg);
}

/* The engine's fast cycle may solve the coarsest level directly (no
 * reference counterpart): B = DH^2 L assembled on the unknowns
 * j*(K-2) + k-1 from the same coefficients the sweeps use, inverted by
 * Gauss-Jordan elimination without pivoting, applied as e = B^{-1}(DH^2 f)
 * (row sums split into column chunks, see orc_coarse_direct). */
static void orc_coarse_inverse(orc_mg *mg) {
  const int l = mg->L - 1;
  const long N = mg->N[l], K = mg->K[l], W = K - 2, n = N * W;
  const long W2 = 2L * (mg->ni + 2) - 1;
  double *B = (double *)calloc((size_t)(n * n), sizeof(double));
  double *X = (double *)calloc((size_t)(n * n), sizeof(double));
  for (long j = 0; j < N; j++)
    for (long k = 1; k <= K - 2; k++) {
      long i1, i2, i3, i4;
      orc_mu_index(l, mg->FC[l], W2, j, k, &i1, &i2, &i3, &i4);
      const double m1 = mg->mu[i1], m2 = mg->mu[i2];
      double m3, m4;
      if (j == 0) {
        m4 = mg->mu[i4];
        m3 = (l == 0) ? mg->mu[i3] : m4;
      } else if (j == N - 1) {
        m3 = mg->mu[i3];
        m4 = (l == 0) ? mg->mu[i4] : m3;
      } else {
        m3 = mg->mu[i3];
        m4 = mg->mu[i4];
      }
      const long jS = (j == 0) ? 1 : j - 1, jN = (j == N - 1) ? N - 2 : j + 1;
      const long row = j * W + (k - 1);
      double *Br = B + row * n;
      if (k - 1 >= 1) Br[j * W + (k - 2)] += m1;
      if (k + 1 <= K - 2) Br[j * W + k] += m2;
      Br[jS * W + (k - 1)] += m3;
      Br[jN * W + (k - 1)] += m4;
      Br[row] -= (m1 + m2 + m3 + m4);
    }
  for (long i = 0; i < n; i++) X[i * n + i] = 1.0;
  for (long p = 0; p < n; p++) {
    double *Ap = B + p * n, *Xp = X + p * n;
    const double piv = Ap[p];
    for (long c = 0; c < n; c++) {
      Ap[c] = Ap[c] / piv;
      Xp[c] = Xp[c] / piv;
    }
    for (long r = 0; r < n; r++) {
      if (r == p) continue;
      double *Ar = B + r * n, *Xr = X + r * n;
      const double f = Ar[p];
      if (f == 0.0) continue;
      for (long c = 0; c < n; c++) {
        Ar[c] = Ar[c] - f * Ap[c];
        Xr[c] = Xr[c] - f * Xp[c];
      }
    }
  }
  /* the engine stores the inverse in single precision (the solve's
   * products and sums stay in double) */
  for (long i = 0; i < n * n; i++) X[i] = (double)(float)X[i];
  free(B);
  free(mg->cinv);
  mg->cinv = X;
  mg->cn = (int)n;
}

static void orc_coarse_direct(orc_mg *mg, double *u, const double *f,
                              int level) {
  const long K = mg->K[level], W = K - 2, n = mg->cn;
  const double DH2 = mg->DH[level] * mg->DH[level];
  double *g = (double *)malloc(sizeof(double) * (size_t)n);
  for (long i = 0; i < n; i++) {
    const long j = i / W, k = i - j * W + 1;
    g[i] = DH2 * f[j * K + k];
  }
  /* columns in Q consecutive chunks of C (Q = min(4, 1024 / n)), each chunk
   * summed in column order from 0, the chunk sums added in chunk order */
  const long Q = (1024 / n) < 4 ? (1024 / n < 1 ? 1 : 1024 / n) : 4;
  const long C = (n + Q - 1) / Q;
#pragma omp parallel for schedule(static) if (n > 256)
  for (long row = 0; row < n; row++) {
    double acc = 0.0;
    for (long q = 0; q < Q; q++) {
      double part = 0.0;
      const long c1 = (q + 1) * C < n ? (q + 1) * C : n;
      for (long col = q * C; col < c1; col++)
        part = part + mg->cinv[row * n + col] * g[col];
      acc = (q == 0) ? part : acc + part;
    }
    const long j = row / W, k = row - j * W + 1;
    u[j * K + k] = acc;
  }
  free(g);
}

void orc_mg_set_cycle(orc_mg *mg, int n_pre, int n_post, int n_coarse,
                      int full_top) {
  mg->n_pre = n_pre;
  mg->n_post = n_post;
  mg->n_coarse = n_coarse;
  mg->full_top = full_top;
  free(mg->cinv);
  mg->cinv = NULL;
  mg->cn = 0;
  if (full_top && n_coarse == 0) orc_coarse_inverse(mg);
}

/* smooth, odeSolver.cpp:741-755: sweeps x (red, black). */
static void orc_smooth(orc_mg *mg, double *u, const double *f, int level,
                       int sweeps) {
  for (int it = 0; it < sweeps; it++) {
    orc_rbgs(u, f, mg->mu, mg->ni, mg->nk, mg->dh, level, 0, mg->relax);
    orc_rbgs(u, f, mg->mu, mg->ni, mg->nk, mg->dh, level, 1, mg->relax);
  }
}

/* vcycle, odeSolver.cpp:1101-1147.  The coarsest level is only smoothed. */
static void orc_vcycle_level(orc_mg *mg, double *u, const double *f,
                             int level) {
  if (level >= mg->L - 1) {
    if (mg->cinv)
      orc_coarse_direct(mg, u, f, level);
    else
      orc_smooth(mg, u, f, level, mg->n_coarse);
    return;
  }
  orc_smooth(mg, u, f, level, mg->n_pre);
  orc_residual(mg->res[level], u, f, mg->mu, mg->ni, mg->nk, mg->dh, level);
  orc_restrict(mg->res2[level], mg->res[level], mg->err2[level], mg->ni,
               mg->nk, level + 1);
  orc_vcycle_level(mg, mg->err2[level], mg->res2[level], level + 1);
  orc_prolong_full(u, mg->err2[level], mg->ni, mg->nk, level + 1,
                   mg->full_top);
  orc_smooth(mg, u, f, level, mg->n_post);
}

void orc_mg_vcycle(orc_mg *mg, double *u, const double *f) {
  orc_vcycle_level(mg, u, f, 0);
}

/* extend_crew, odeSolver.cpp:612-650: linear extrapolation of the ghost
 * columns, then mirrored ghost rows. */
void orc_extend_crew(double *U, int ni, int nk) {
  const long nx = ni + 2, nz = nk + 2;
  for (long k = 0; k < nz; k++) {
    long index1 = 1 + k * nx, index2 = ni + k * nx;
    U[index1 - 1] = 2.0 * U[index1] - U[index1 + 1];
    U[index2 + 1] = 2.0 * U[index2] - U[index2 - 1];
  }
  for (long i = 0; i < nx; i++) {
    long index1 = i + 1 * nx, index2 = i + (long)nk * nx;
    U[index1 - nx] = U[index1 + nx];
    U[index2 + nx] = U[index2 - nx];
  }
}

/* ode_solve (odeSolver.cpp:605-609) = multigrid (652-720) + extend_crew.
 * Pack U's interior into the compact u, iterate V-cycles while the level-0
 * residual norm exceeds the threshold, unpack.  The initial guess is the
 * previous solve's U (Q8). */
int orc_ode_solve(orc_mg *mg, double *U, int tol_mode, double tol,
                  int max_vcycles, int *vcycles, double *res0, double *res) {
  const int ni = mg->ni, nk = mg->nk;
  const long nx = ni + 2, K = ni;
  const long n0 = (long)ni * nk;
#pragma omp parallel for schedule(static) if (n0 > ORC_OMP_MIN)
  for (long j = 1; j <= nk; j++)
    for (long i = 1; i <= ni; i++) mg->u[(i - 1) + (j - 1) * K] = U[i + j * nx];

  double thr = tol;
  double r0 = 0.0;
  if (tol_mode == ORC_TOL_REL) {
    /* boundary-data residual: interior zero, Dirichlet columns as set */
    memset(mg->tmp, 0, sizeof(double) * (size_t)n0);
    for (long j = 0; j < nk; j++) {
      mg->tmp[0 + j * K] = mg->u[0 + j * K];
      mg->tmp[(K - 1) + j * K] = mg->u[(K - 1) + j * K];
    }
    orc_residual(mg->delta, mg->tmp, mg->f, mg->mu, ni, nk, mg->dh, 0);
    r0 = orc_nrm2(mg->delta, n0);
    thr = tol * r0;
  }
  orc_residual(mg->delta, mg->u, mg->f, mg->mu, ni, nk, mg->dh, 0);
  double delta = orc_nrm2(mg->delta, n0);
  if (tol_mode != ORC_TOL_REL) r0 = delta;
  int it = 0, status = ORC_OK;
  while (delta > thr) {
    if (max_vcycles > 0 && it >= max_vcycles) {
      status = ORC_ERR_MG_NOT_CONVERGED;
      break;
    }
    orc_mg_vcycle(mg, mg->u, mg->f);
    orc_residual(mg->delta, mg->u, mg->f, mg->mu, ni, nk, mg->dh, 0);
    delta = orc_nrm2(mg->delta, n0);
    it++;
  }
#pragma omp parallel for schedule(static) if (n0 > ORC_OMP_MIN)
  for (long j = 1; j <= nk; j++)
    for (long i = 1; i <= ni; i++) U[i + j * nx] = mg->u[(i - 1) + (j - 1) * K];
  orc_extend_crew(U, ni, nk);
  if (vcycles) *vcycles = it;
  if (res0) *res0 = r0;
  if (res) *res = delta;
  return status;
}

/* ------------------------------------------------------------------------ */
/* Elementary functions of the fault update.                                 */
/*                                                                           */
/* The reference calls the platform libm (glibc) for exp, sinh, cosh and     */
/* asinh in NRsearch and cal_dpsi (fault.cpp:184-228, odeSolver.cpp:1198).  */
/* The GPU cannot reproduce glibc bit for bit, so the engine evaluates these */
/* with its own algorithms written in +,-,*,/, sqrt, floor, frexp and ldexp  */
/* (Cody-Waite exp with a Taylor polynomial, Taylor sinh for |x| < 1/2,      */
/* atanh-series log, Goldberg's log1p).  The oracle carries the same         */
/* algorithms, selected by default, so the two sides agree bit for bit; with */
/* orc_set_libm(1) the oracle uses glibc instead, which measures how far the */
/* choice of libm moves the results.                                         */
/* ------------------------------------------------------------------------ */

static const double Q_LN2_HI = 6.93147180369123816490e-01;
static const double Q_LN2_LO = 1.90821492927058770002e-10;
static const double Q_INV_LN2 = 1.44269504088896338700e+00;
static const double Q_LN2 = 6.93147180559945309417e-01;

static double q_exp(double x) {
  if (x != x) return x;
  if (x > 709.782712893383973096) return HUGE_VAL;
  if (x < -745.133219101941108420) return 0.0;
  const double kd = floor(x * Q_INV_LN2 + 0.5);
  const int k = (int)kd;
  const double r = (x - kd * Q_LN2_HI) - kd * Q_LN2_LO;
  double p = 1.0 / 6227020800.0;
  p = p * r + 1.0 / 479001600.0;
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 0.5;
  p = p * r + 1.0;
  const double y = 1.0 + p * r;
  if (k > 1020) return ldexp(ldexp(y, k - 1020), 1020);
  if (k < -1020) return ldexp(ldexp(y, k + 1020), -1020);
  return ldexp(y, k);
}

static double q_sinh(double x) {
  if (x != x) return x;
  const double ax = fabs(x);
  double r;
  if (ax < 0.5) {
    const double x2 = ax * ax;
    double p = 1.0 / 355687428096000.0;
    p = p * x2 + 1.0 / 1307674368000.0;
    p = p * x2 + 1.0 / 6227020800.0;
    p = p * x2 + 1.0 / 39916800.0;
    p = p * x2 + 1.0 / 362880.0;
    p = p * x2 + 1.0 / 5040.0;
    p = p * x2 + 1.0 / 120.0;
    p = p * x2 + 1.0 / 6.0;
    r = ax + ax * (x2 * p);
  } else if (ax < 22.0) {
    const double e = q_exp(ax);
    r = 0.5 * (e - 1.0 / e);
  } else if (ax < 709.0) {
    r = 0.5 * q_exp(ax);
  } else {
    const double e = q_exp(0.5 * ax);
    r = (0.5 * e) * e;
  }
  return x < 0.0 ? -r : r;
}

static double q_cosh(double x) {
  if (x != x) return x;
  const double ax = fabs(x);
  if (ax < 22.0) {
    const double e = q_exp(ax);
    return 0.5 * (e + 1.0 / e);
  }
  if (ax < 709.0) return 0.5 * q_exp(ax);
  const double e = q_exp(0.5 * ax);
  return (0.5 * e) * e;
}

static double q_log(double x) {
  if (x != x) return x;
  if (x < 0.0) return (x - x) / (x - x);
  if (x == 0.0) return -HUGE_VAL;
  if (x > 1.79769313486231570815e+308) return x;
  int e;
  double m = frexp(x, &e);
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e = e - 1;
  }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double s2 = s * s;
  double p = 1.0 / 25.0;
  p = p * s2 + 1.0 / 23.0;
  p = p * s2 + 1.0 / 21.0;
  p = p * s2 + 1.0 / 19.0;
  p = p * s2 + 1.0 / 17.0;
  p = p * s2 + 1.0 / 15.0;
  p = p * s2 + 1.0 / 13.0;
  p = p * s2 + 1.0 / 11.0;
  p = p * s2 + 1.0 / 9.0;
  p = p * s2 + 1.0 / 7.0;
  p = p * s2 + 1.0 / 5.0;
  p = p * s2 + 1.0 / 3.0;
  const double lm = 2.0 * s + 2.0 * s * (s2 * p);
  const double ed = (double)e;
  return ed * Q_LN2_HI + (ed * Q_LN2_LO + lm);
}

static double q_log1p(double x) {
  const double u = 1.0 + x;
  if (u == 1.0) return x;
  return q_log(u) * (x / (u - 1.0));
}

static double q_asinh(double x) {
  if (x != x) return x;
  const double ax = fabs(x);
  double r;
  if (ax < 3.7252902984619140625e-09) return x;
  if (ax > 268435456.0) {
    r = q_log(ax) + Q_LN2;
  } else if (ax > 2.0) {
    r = q_log(2.0 * ax + 1.0 / (sqrt(ax * ax + 1.0) + ax));
  } else {
    const double t = ax * ax;
    r = q_log1p(ax + t / (1.0 + sqrt(1.0 + t)));
  }
  return x < 0.0 ? -r : r;
}

static double (*f_exp)(double) = q_exp;
static double (*f_sinh)(double) = q_sinh;
static double (*f_cosh)(double) = q_cosh;
static double (*f_asinh)(double) = q_asinh;

void orc_set_libm(int use_glibc) {
  f_exp = use_glibc ? exp : q_exp;
  f_sinh = use_glibc ? sinh : q_sinh;
  f_cosh = use_glibc ? cosh : q_cosh;
  f_asinh = use_glibc ? asinh : q_asinh;
}

double orc_elem(int fn, double x) {
  switch (fn) {
    case 0: return q_exp(x);
    case 1: return q_sinh(x);
    case 2: return q_cosh(x);
    case 3: return q_log(x);
    case 4: return q_log1p(x);
    case 5: return q_asinh(x);
    default: return 0.0 / 0.0;
  }
}

/* ------------------------------------------------------------------------ */
/* Fault physics                                                             */
/* ------------------------------------------------------------------------ */

/* setBoundary, odeSolver.cpp:1150-1188. */
void orc_set_boundary(double *U, int ni, int nk, const double *slip,
                      const double *slip_shift, double Vini, double t,
                      double ratio, double dt) {
  const long nx = ni + 2, nz = nk + 2;
  for (long k = 1; k <= nk; k++) U[k * nx + 1] = 0.5 * slip[k];
  for (long k = 0; k < nz; k++)
    U[k * nx + ni] = 0.50 * Vini * (t + ratio * dt) + slip_shift[k];
  for (long k = 0; k < nx; k++) U[0 * nx + k] = U[2 * nx + k];
  for (long k = 0; k < nx; k++) U[(nk + 1) * nx + k] = U[(nk - 1) * nx + k];
}

/* eval_stress_on_fault + partial1st, fault.cpp:139-164: one-sided 4-point
 * x-derivative at the fault column times nodal mu, x1e3 (GPa -> MPa). */
void orc_eval_stress(const double *U, const double *mu, const double *tau_n0,
                     int ni, int nk, double dh, double *Ts, double *Tn) {
  const long nx = ni + 2;
  for (long k = 1; k <= nk; k++) {
    long index = 1 + k * nx;
    double du = (-24.0 / (17.0 * dh)) * U[index] +
                (59.0 / (34.0 * dh)) * U[index + 1] +
                (-4.0 / (17.0 * dh)) * U[index + 2] +
                (-3.0 / (34.0 * dh)) * U[index + 3];
    Ts[k] = mu[index] * du * 1.0e3;
    Tn[k] = tau_n0[k];
  }
}

/* NRsearch, fault.cpp:169-232: Newton on x = asinh(V/(2V0) e^{psi/a}) for
 * tau = a|sigma| x + 2 eta V0 e^{-psi/a} sinh(x).  V is the previous slip
 * rate on entry (the Newton start) and the new slip rate on exit.  A NaN
 * result is retried once and then reported (the reference calls exit(0)). */
int orc_nr_search(int nk, const double *Ts, const double *Tn,
                  const double *psi, const double *a, const double *eta,
                  double V0, double *V, int *iters) {
  for (long k = 1; k <= nk; k++) {
    int flag_nan = 0, iter;
    double var_old, var_new = 0.0, fa, fb, fy, df, d, err;
  retry:
    iter = 0;
    var_old = 0.5 * V[k] / V0 * f_exp(psi[k] / a[k]);
    var_old = f_asinh(var_old);
    fa = a[k] * fabs(Tn[k]);
    fb = Ts[k];
    const int iter_max = 100000;
    while (iter < iter_max) {
      fy = 2.0 * eta[k] * V0 * f_exp(-psi[k] / a[k]) * f_sinh(var_old) +
           fa * var_old - fb;
      df = 2.0 * eta[k] * V0 * f_exp(-psi[k] / a[k]) * f_cosh(var_old) + fa;
      d = -fy / (df + 1e-100);
      var_new = var_old + d;
      err = fabs(d) / (fabs(var_old) + 1e-100);
      iter = iter + 1;
      var_old = var_new;
      if (err < 1e-10) break;
    }
    double friction_update = a[k] * var_new;
    if (isnan(friction_update)) {
      if (flag_nan) return ORC_ERR_NAN_FRICTION;
      flag_nan = 1;
      goto retry;
    }
    if (iters) iters[k] = iter;
    V[k] = 2.0 * V0 * f_exp(-psi[k] / a[k]);
    V[k] = V[k] * f_sinh(var_new);
  }
  return ORC_OK;
}

/* cal_dpsi, odeSolver.cpp:1190-1206: aging law in psi form; zero where
 * b <= 1e-3 or the exponential overflows. */
void orc_cal_dpsi(int nk, const double *psi, const double *V, const double *b,
                  double f0, double V0, double L, double *dpsi) {
  for (long k = 1; k <= nk; k++) {
    double A = f_exp((f0 - psi[k]) / b[k]);
    dpsi[k] = 0.0;
    if (!isinf(A) && b[k] > 1e-3) dpsi[k] = (b[k] * V0 / L) * (A - V[k] / V0);
  }
}

/* computeError, odeSolver.cpp:1211-1238. */
double orc_compute_error(int nk, const double *slip3, const double *psi3,
                         const double *slip4, const double *psi4) {
  double err1 = 0.0, err2 = 0.0;
  for (long k = 1; k <= nk; k++) {
    double t1 = slip3[k] - slip4[k];
    double t2 = psi3[k] - psi4[k];
    err1 = err1 + t1 * t1;
    err2 = err2 + t2 * t2;
  }
  err1 = sqrt(err1);
  err2 = sqrt(err2);
  return err1 / (sqrt((double)nk) * 1.0) + err2 / (sqrt((double)nk) * 1.0);
}

/* computeStepSize, odeSolver.cpp:1241-1271: elementary controller for the
 * first three steps, PID afterwards; clamp to [mindt, min(5 dt, maxdt)]. */
double orc_compute_step_size(int it, double dt, double mindt, double maxdt,
                             double totErr, double totTol,
                             const double *errA) {
  const double ord = 4.0, kappa = 0.9;
  const double alpha = 0.49 / ord, beta = 0.34 / ord, gamma = 0.1 / ord;
  double stepRatio, deltaT;
  if (it < 4)
    stepRatio = kappa * pow(totTol / totErr, 1.0 / (1.0 + ord));
  else
    stepRatio = kappa * pow(totTol / totErr, alpha) *
                pow(errA[0] / totTol, beta) * pow(totTol / errA[1], gamma);
  deltaT = stepRatio * dt;
  deltaT = OMIN(dt * 5.0, deltaT);
  deltaT = OMIN(maxdt, deltaT);
  deltaT = OMAX(mindt, deltaT);
  return deltaT;
}

/* ------------------------------------------------------------------------ */
/* RK43 driver, odeSolver.cpp:1277-1674                                      */
/* ------------------------------------------------------------------------ */

struct orc_sim {
  orc_params p;
  long nx, nz;
  double *mu, *a, *b, *eta, *tau_n0, *slip_shift;
  double *U, *slip, *psi, *Ts, *Tn;
  double *ks[7], *kp[7];   /* stage states k1..k6 (index 1..6) */
  double *fV[7], *fd[7];   /* stage rates f1..f6 */
  double *y3s, *y3p, *y4s, *y4p;
  double t, dt, totErr, errA[2];
  int it, attempts;
  long vcycles, solves;
  orc_mg *mg;
  /* ring of the last 8 stage calls (t, dt, ratio, remote and fault column
   * values of row 1 right after setBoundary); log_n calls so far */
  double log[8][5];
  long log_n;
};

/* Planted mistakes for the tests that show the RK43 checks detect them: the
 * tests build this file with -DORC_MUTANT=n into a scratch directory.
 * 1 swaps a42 and a43 in stage 4's sums, 2 hands stage 4 the boundary time
 * of stage 3, 3 commits the embedded third-order solution.  0 (the default)
 * is the restatement itself. */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

/* Butcher tableau, odeSolver.cpp:1297-1334. */
static const double c2 = 1. / 2., c3 = 83. / 250., c4 = 31. / 50.,
                    c5 = 17. / 20., c6 = 1.;
static const double b1 = 82889. / 524892., b3 = 15625. / 83664.,
                    b4 = 69875. / 102672., b5 = -2260. / 8211., b6 = 1. / 4.;
static const double hb1 = 4586570599. / 29645900160.,
                    hb3 = 178811875. / 945068544.,
                    hb4 = 814220225. / 1159782912.,
                    hb5 = -3700637. / 11593932., hb6 = 61727. / 225920.;
static const double a21 = 1. / 2.;
static const double a31 = 13861. / 62500., a32 = 6889. / 62500.;
static const double a41 = -116923316275. / 2393684061468.,
                    a42 = -2731218467317. / 15368042101831.,
                    a43 = 9408046702089. / 11113171139209.;
static const double a51 = -451086348788. / 2902428689909.,
                    a52 = -2682348792572. / 7519795681897.,
                    a53 = 12662868775082. / 11960479115383.,
                    a54 = 3355817975965. / 11060851509271.;
static const double a61 = 647845179188. / 3216320057751.,
                    a62 = 73281519250. / 8382639484533.,
                    a63 = 552539513391. / 3454668386233.,
                    a64 = 3354512671639. / 8306763924573.,
                    a65 = 4040. / 17871.;

/* The tableau as used above: out[0..5] c1..c6, out[6..11] b1..b6,
 * out[12..17] hb1..hb6, out[18..53] a[i][j] (6x6, row-major, i,j from 1). */
void orc_tableau(double *out) {
  const double c[6] = {0.0, c2, c3, c4, c5, c6};
  const double b[6] = {b1, 0.0, b3, b4, b5, b6};
  const double hb[6] = {hb1, 0.0, hb3, hb4, hb5, hb6};
  const double a[6][6] = {{0, 0, 0, 0, 0, 0},
                          {a21, 0, 0, 0, 0, 0},
                          {a31, a32, 0, 0, 0, 0},
                          {a41, a42, a43, 0, 0, 0},
                          {a51, a52, a53, a54, 0, 0},
                          {a61, a62, a63, a64, a65, 0}};
  for (int i = 0; i < 6; i++) {
    out[i] = c[i];
    out[6 + i] = b[i];
    out[12 + i] = hb[i];
    for (int j = 0; j < 6; j++) out[18 + 6 * i + j] = a[i][j];
  }
}

static double *orc_dup(const double *x, size_t n) {
  double *y = (double *)malloc(n * sizeof(double));
  memcpy(y, x, n * sizeof(double));
  return y;
}

orc_sim *orc_sim_create(const orc_params *p, const double *mu,
                        const double *mu_equi, const double *a,
                        const double *b, const double *eta,
                        const double *tau_n0, const double *U0) {
  orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
  s->p = *p;
  s->nx = p->ni + 2;
  s->nz = p->nk + 2;
  const size_t ng = (size_t)(s->nx * s->nz), nz = (size_t)s->nz;
  s->mu = orc_dup(mu, ng);
  s->a = orc_dup(a, nz);
  s->b = orc_dup(b, nz);
  s->eta = orc_dup(eta, nz);
  s->tau_n0 = orc_dup(tau_n0, nz);
  s->U = orc_dup(U0, ng);
  s->slip = (double *)calloc(nz, sizeof(double));
  s->psi = (double *)calloc(nz, sizeof(double));
  s->Ts = (double *)calloc(nz, sizeof(double));
  s->Tn = (double *)calloc(nz, sizeof(double));
  s->slip_shift = (double *)calloc(nz, sizeof(double));
  for (int i = 1; i <= 6; i++) {
    s->ks[i] = (double *)calloc(nz, sizeof(double));
    s->kp[i] = (double *)calloc(nz, sizeof(double));
    s->fV[i] = (double *)calloc(nz, sizeof(double));
    s->fd[i] = (double *)calloc(nz, sizeof(double));
  }
  s->y3s = (double *)calloc(nz, sizeof(double));
  s->y3p = (double *)calloc(nz, sizeof(double));
  s->y4s = (double *)calloc(nz, sizeof(double));
  s->y4p = (double *)calloc(nz, sizeof(double));
  /* fault_init, fault.cpp:116-133 */
  for (long k = 1; k <= p->nk; k++) {
    s->psi[k] = log(p->V0 / p->Vini);
    s->psi[k] = p->f0 + s->b[k] * s->psi[k];
  }
  for (long k = 1; k <= p->nk; k++) s->slip[k] = 2.0 * s->U[k * s->nx + 1];
  for (long k = 0; k < s->nz; k++) s->slip_shift[k] = s->U[k * s->nx + p->ni];
  s->dt = p->ini_dt; /* odeSolver.cpp:1363 */
  s->t = 0.0;
  s->mg = orc_mg_create(p->ni, p->nk, p->dh, p->n_levels, p->n_cycles,
                        p->relax, mu_equi);
  if (p->cycle == 1)
    orc_mg_set_cycle(s->mg, p->n_pre, p->n_post, p->n_coarse, 1);
  return s;
}

void orc_sim_destroy(orc_sim *s) {
  if (!s) return;
  orc_mg_destroy(s->mg);
  free(s->mu); free(s->a); free(s->b); free(s->eta); free(s->tau_n0);
  free(s->U); free(s->slip); free(s->psi); free(s->Ts); free(s->Tn);
  free(s->slip_shift);
  for (int i = 1; i <= 6; i++) {
    free(s->ks[i]); free(s->kp[i]); free(s->fV[i]); free(s->fd[i]);
  }
  free(s->y3s); free(s->y3p); free(s->y4s); free(s->y4p);
  free(s);
}

/* one stage: setBoundary -> ode_solve -> eval_stress -> NRsearch -> cal_dpsi
 * (e.g. odeSolver.cpp:1496-1500) */
static int orc_stage(orc_sim *s, const double *ys, const double *yp,
                     double ratio, double *fV, double *fd) {
  const orc_params *p = &s->p;
  int vc = 0, st;
  orc_set_boundary(s->U, p->ni, p->nk, ys, s->slip_shift, p->Vini, s->t,
                   ratio, s->dt);
  {
    double *r = s->log[s->log_n % 8];
    r[0] = s->t;
    r[1] = s->dt;
    r[2] = ratio;
    r[3] = s->U[1 * s->nx + p->ni];
    r[4] = s->U[1 * s->nx + 1];
    s->log_n++;
  }
  st = orc_ode_solve(s->mg, s->U, p->tol_mode, p->tol, p->max_vcycles, &vc,
                     NULL, NULL);
  s->vcycles += vc;
  s->solves += 1;
  if (st != ORC_OK) return st;
  orc_eval_stress(s->U, s->mu, s->tau_n0, p->ni, p->nk, p->dh, s->Ts, s->Tn);
  st = orc_nr_search(p->nk, s->Ts, s->Tn, yp, s->a, s->eta, p->V0, fV, NULL);
  if (st != ORC_OK) return st;
  orc_cal_dpsi(p->nk, yp, fV, s->b, p->f0, p->V0, p->L, fd);
  return ORC_OK;
}

/* Restart from a saved state: t, dt, step count, controller history and
 * the committed (U, slip, psi); stage 1 is recomputed as after an accepted
 * step (odeSolver.cpp:1615-1619). */
int orc_sim_set_state(orc_sim *s, double t, double dt, int it, double errA0,
                      double errA1, const double *U, const double *slip,
                      const double *psi) {
  const size_t ng = (size_t)(s->nx * s->nz), nz = (size_t)s->nz;
  s->t = t;
  s->dt = dt;
  s->it = it;
  s->errA[0] = errA0;
  s->errA[1] = errA1;
  memcpy(s->U, U, ng * sizeof(double));
  memcpy(s->slip, slip, nz * sizeof(double));
  memcpy(s->psi, psi, nz * sizeof(double));
  for (int i = 1; i <= 6; i++) {
    memset(s->fV[i], 0, nz * sizeof(double));
    memset(s->fd[i], 0, nz * sizeof(double));
  }
  return orc_sim_init(s);
}

int orc_sim_init(orc_sim *s) {
  for (long k = 1; k <= s->p.nk; k++) { /* 1428-1432 */
    s->ks[1][k] = s->slip[k];
    s->kp[1][k] = s->psi[k];
  }
  return orc_stage(s, s->ks[1], s->kp[1], 0.0, s->fV[1], s->fd[1]);
}

int orc_sim_step(orc_sim *s) {
  const orc_params *p = &s->p;
  const long nk = p->nk;
  const size_t nzb = sizeof(double) * (size_t)s->nz;
  double *const *fV = s->fV, *const *fd = s->fd;
  int st;
  s->it = s->it + 1;
  int attemptCount = 0;
  while (attemptCount < 100) { /* 1466-1604 */
    attemptCount = attemptCount + 1;
    if (s->t + s->dt > p->tmax) s->dt = p->tmax - s->t;
    for (int i = 2; i <= 6; i++) {
      memset(s->ks[i], 0, nzb); memset(s->kp[i], 0, nzb);
      memset(s->fV[i], 0, nzb); memset(s->fd[i], 0, nzb);
    }
    memset(s->y3s, 0, nzb); memset(s->y3p, 0, nzb);
    memset(s->y4s, 0, nzb); memset(s->y4p, 0, nzb);
    const double dt = s->dt;
    double *k2s = s->ks[2], *k2p = s->kp[2];
    for (long k = 1; k <= nk; k++) {
      k2s[k] = s->slip[k] + a21 * dt * fV[1][k];
      k2p[k] = s->psi[k] + a21 * dt * fd[1][k];
    }
    if ((st = orc_stage(s, k2s, k2p, c2, fV[2], fd[2])) != ORC_OK) return st;
    double *k3s = s->ks[3], *k3p = s->kp[3];
    for (long k = 1; k <= nk; k++) {
      k3s[k] = s->slip[k] + a31 * dt * fV[1][k];
      k3s[k] = k3s[k] + a32 * dt * fV[2][k];
      k3p[k] = s->psi[k] + a31 * dt * fd[1][k];
      k3p[k] = k3p[k] + a32 * dt * fd[2][k];
    }
    if ((st = orc_stage(s, k3s, k3p, c3, fV[3], fd[3])) != ORC_OK) return st;
    double *k4s = s->ks[4], *k4p = s->kp[4];
    for (long k = 1; k <= nk; k++) {
      k4s[k] = s->slip[k] + a41 * dt * fV[1][k];
      k4s[k] = k4s[k] + (ORC_MUTANT == 1 ? a43 : a42) * dt * fV[2][k];
      k4s[k] = k4s[k] + (ORC_MUTANT == 1 ? a42 : a43) * dt * fV[3][k];
      k4p[k] = s->psi[k] + a41 * dt * fd[1][k];
      k4p[k] = k4p[k] + (ORC_MUTANT == 1 ? a43 : a42) * dt * fd[2][k];
      k4p[k] = k4p[k] + (ORC_MUTANT == 1 ? a42 : a43) * dt * fd[3][k];
    }
    if ((st = orc_stage(s, k4s, k4p, ORC_MUTANT == 2 ? c3 : c4, fV[4], fd[4])) !=
        ORC_OK)
      return st;
    double *k5s = s->ks[5], *k5p = s->kp[5];
    for (long k = 1; k <= nk; k++) {
      k5s[k] = s->slip[k] + a51 * dt * fV[1][k];
      k5s[k] = k5s[k] + a52 * dt * fV[2][k];
      k5s[k] = k5s[k] + a53 * dt * fV[3][k];
      k5s[k] = k5s[k] + a54 * dt * fV[4][k];
      k5p[k] = s->psi[k] + a51 * dt * fd[1][k];
      k5p[k] = k5p[k] + a52 * dt * fd[2][k];
      k5p[k] = k5p[k] + a53 * dt * fd[3][k];
      k5p[k] = k5p[k] + a54 * dt * fd[4][k];
    }
    if ((st = orc_stage(s, k5s, k5p, c5, fV[5], fd[5])) != ORC_OK) return st;
    double *k6s = s->ks[6], *k6p = s->kp[6];
    for (long k = 1; k <= nk; k++) {
      k6s[k] = s->slip[k] + a61 * dt * fV[1][k];
      k6s[k] = k6s[k] + a62 * dt * fV[2][k];
      k6s[k] = k6s[k] + a63 * dt * fV[3][k];
      k6s[k] = k6s[k] + a64 * dt * fV[4][k];
      k6s[k] = k6s[k] + a65 * dt * fV[5][k];
      k6p[k] = s->psi[k] + a61 * dt * fd[1][k];
      k6p[k] = k6p[k] + a62 * dt * fd[2][k];
      k6p[k] = k6p[k] + a63 * dt * fd[3][k];
      k6p[k] = k6p[k] + a64 * dt * fd[4][k];
      k6p[k] = k6p[k] + a65 * dt * fd[5][k];
    }
    if ((st = orc_stage(s, k6s, k6p, c6, fV[6], fd[6])) != ORC_OK) return st;
    for (long k = 1; k <= nk; k++) { /* 1569-1593 */
      s->y3s[k] = s->slip[k] + hb1 * dt * fV[1][k];
      s->y3s[k] = s->y3s[k] + hb3 * dt * fV[3][k];
      s->y3s[k] = s->y3s[k] + hb4 * dt * fV[4][k];
      s->y3s[k] = s->y3s[k] + hb5 * dt * fV[5][k];
      s->y3s[k] = s->y3s[k] + hb6 * dt * fV[6][k];
      s->y3p[k] = s->psi[k] + hb1 * dt * fd[1][k];
      s->y3p[k] = s->y3p[k] + hb3 * dt * fd[3][k];
      s->y3p[k] = s->y3p[k] + hb4 * dt * fd[4][k];
      s->y3p[k] = s->y3p[k] + hb5 * dt * fd[5][k];
      s->y3p[k] = s->y3p[k] + hb6 * dt * fd[6][k];

      s->y4s[k] = s->slip[k] + b1 * dt * fV[1][k];
      s->y4s[k] = s->y4s[k] + b3 * dt * fV[3][k];
      s->y4s[k] = s->y4s[k] + b4 * dt * fV[4][k];
      s->y4s[k] = s->y4s[k] + b5 * dt * fV[5][k];
      s->y4s[k] = s->y4s[k] + b6 * dt * fV[6][k];
      s->y4p[k] = s->psi[k] + b1 * dt * fd[1][k];
      s->y4p[k] = s->y4p[k] + b3 * dt * fd[3][k];
      s->y4p[k] = s->y4p[k] + b4 * dt * fd[4][k];
      s->y4p[k] = s->y4p[k] + b5 * dt * fd[5][k];
      s->y4p[k] = s->y4p[k] + b6 * dt * fd[6][k];
    }
    s->totErr = orc_compute_error(p->nk, s->y3s, s->y3p, s->y4s, s->y4p);
    if (s->totErr < p->totTol) break;
    s->dt = orc_compute_step_size(s->it, s->dt, p->mindt, p->maxdt, s->totErr,
                                  p->totTol, s->errA);
    if (p->mindt == s->dt) break;
  }
  s->attempts = attemptCount;
  s->t = s->t + s->dt; /* 1605-1614 */
  {
    const double *ns = ORC_MUTANT == 3 ? s->y3s : s->y4s;
    const double *np = ORC_MUTANT == 3 ? s->y3p : s->y4p;
    for (long k = 1; k <= nk; k++) {
      s->slip[k] = ns[k];
      s->ks[1][k] = ns[k];
      fV[1][k] = 0.0;
      s->psi[k] = np[k];
      s->kp[1][k] = np[k];
      fd[1][k] = 0.0;
    }
  }
  if ((st = orc_stage(s, s->ks[1], s->kp[1], 0.0, fV[1], fd[1])) != ORC_OK)
    return st;
  double newDeltaT; /* 1621-1632 */
  if (s->totErr != 0.0)
    newDeltaT = orc_compute_step_size(s->it, s->dt, p->mindt, p->maxdt,
                                      s->totErr, p->totTol, s->errA);
  else
    newDeltaT = s->dt;
  s->errA[1] = s->errA[0];
  s->errA[0] = s->totErr;
  s->dt = newDeltaT;
  return ORC_OK;
}

void orc_sim_set_dt(orc_sim *s, double dt) { s->dt = dt; }

/* Stage data of the last attempt of the last step: the stage rates fV, fd
 * and states ks, kp of stages 2..6 (nz each, at offset (i-1)*nz for stage i;
 * stage 1's entries are zero here, its values being those before the step)
 * and the embedded (y3) and propagated (y4) solutions. */
void orc_sim_last_attempt(const orc_sim *s, double *fV, double *fd, double *ks,
                          double *kp, double *y3s, double *y3p, double *y4s,
                          double *y4p) {
  const size_t nz = (size_t)s->nz, nb = nz * sizeof(double);
  for (int i = 1; i <= 6; i++) {
    const size_t o = (size_t)(i - 1) * nz;
    if (i == 1) {
      memset(fV + o, 0, nb); memset(fd + o, 0, nb);
      memset(ks + o, 0, nb); memset(kp + o, 0, nb);
      continue;
    }
    memcpy(fV + o, s->fV[i], nb); memcpy(fd + o, s->fd[i], nb);
    memcpy(ks + o, s->ks[i], nb); memcpy(kp + o, s->kp[i], nb);
  }
  memcpy(y3s, s->y3s, nb); memcpy(y3p, s->y3p, nb);
  memcpy(y4s, s->y4s, nb); memcpy(y4p, s->y4p, nb);
}

/* The last min(8, calls) stage calls, oldest first, 5 doubles each:
 * t, dt, ratio, U(row 1, remote column), U(row 1, fault column) right after
 * setBoundary.  Returns the number of records. */
int orc_sim_stage_log(const orc_sim *s, double *out) {
  const long n = s->log_n < 8 ? s->log_n : 8;
  for (long q = 0; q < n; q++) {
    const double *r = s->log[(s->log_n - n + q) % 8];
    for (int i = 0; i < 5; i++) out[5 * q + i] = r[i];
  }
  return (int)n;
}

/* maxval, fault.cpp:22-52 */
static double orc_maxval(const double *x, long nk) {
  double m = x[1];
  for (long k = 2; k <= nk; k++)
    if (m < x[k]) m = x[k];
  return m;
}

void orc_sim_scalars(const orc_sim *s, double *o) {
  o[0] = s->t;
  o[1] = s->dt;
  o[2] = s->it;
  o[3] = s->errA[0];
  o[4] = s->errA[1];
  o[5] = s->totErr;
  o[6] = s->attempts;
  o[7] = (double)s->vcycles;
  o[8] = (double)s->solves;
  o[9] = orc_maxval(s->fV[1], s->p.nk);
  o[10] = orc_maxval(s->Ts, s->p.nk);
}

void orc_sim_arrays(const orc_sim *s, double *U, double *slip, double *psi,
                    double *V, double *dpsi, double *Ts) {
  const size_t ng = (size_t)(s->nx * s->nz), nz = (size_t)s->nz;
  if (U) memcpy(U, s->U, ng * sizeof(double));
  if (slip) memcpy(slip, s->slip, nz * sizeof(double));
  if (psi) memcpy(psi, s->psi, nz * sizeof(double));
  if (V) memcpy(V, s->fV[1], nz * sizeof(double));
  if (dpsi) memcpy(dpsi, s->fd[1], nz * sizeof(double));
  if (Ts) memcpy(Ts, s->Ts, nz * sizeof(double));
}
