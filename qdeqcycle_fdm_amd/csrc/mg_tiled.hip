// mg_tiled.hip -- the fast V-cycle: temporally blocked LDS tile kernels.
//
// One V-cycle of the fast cycle costs two launches per level above the coarse
// tail plus one launch for the whole tail:
//   k_down  (level l): n_pre red/black sweeps, residual, full-weighting
//           restriction to level l+1, all on one LDS tile with a halo of
//           2 n_pre + 2 points, so the ring of redundant updates makes every
//           interior value exactly what global sweeps would produce;
//   k_tail  (levels with N <= TAIL_MAX_N): one workgroup runs the rest of the
//           V-cycle in LDS with workgroup barriers only;
//   k_up    (level l): bilinear prolongation + correction (folded into the
//           tile load), n_post sweeps, and on level 0 either the
//           residual-norm partials (FP64 solve) or the FP64 update u += e
//           (FP32 correction of iterative refinement).
// Every loop over a tile window deals its points to all threads of the
// workgroup in turn (FlatWalk).
// Every update, residual, restriction and prolongation is evaluated with the
// same expressions as the one-launch-per-half-sweep kernels of mg.hip, so the
// tiled cycle reproduces them bit for bit.  The input and output u of a tile
// kernel are different buffers (tiles read their neighbours' interiors as
// halo).
#include <cmath>

#include <algorithm>
#include <utility>
#include <vector>

#include "qdeq_internal.h"

namespace qdeq {

constexpr int TX_THREADS = 64;   // one wavefront across a tile row
constexpr int TY_THREADS = 16;   // 16 rows of lanes: 1024 threads per tile
constexpr int TILE_THREADS = TX_THREADS * TY_THREADS;
constexpr int TAIL_THREADS = 1024;
constexpr int TAIL_MAX_N = 33;  // levels with N, K <= this run in the tail
// the tail loads a level's arrays in at most two passes of its threads
static_assert((TAIL_MAX_N + 1) * TAIL_MAX_N <= 2 * TAIL_THREADS, "tail load passes");

template <typename T, bool EXACT>
__device__ inline T rbgs_update(T m1, T m2, T m3, T m4, T uW, T uE, T uS,
                                T uN, T uC, T fDH2, bool has_f, T relax) {
  T num = m1 * uW + m2 * uE + m3 * uS + m4 * uN;
  if (has_f) num = num - fDH2;
  if (EXACT) return (relax / (m1 + m2 + m3 + m4)) * num + (T(1) - relax) * uC;
  const T unew = num / (m1 + m2 + m3 + m4);
  return (relax == T(1)) ? unew : uC + relax * (unew - uC);
}

// Shared layout of a tile: rows [0, R), columns [0, C), row stride LS.
// su, sf, scx hold R rows; scz holds R+1 rows (south faces; row lr+1 is the
// north face of row lr).  Global row of local row lr is j0 + lr.
template <typename T>
struct Tile {
  T *su, *sf, *scx, *scz;
  int R, C, LS, j0, k0;
};

// Points (r, c) of a rows x cols window dealt to the workgroup's threads in
// turn, row by row: thread tid takes points tid, tid + nthr, ...  Every lane
// has a point even where a row is shorter than a wavefront (a 66-point tile
// row would leave a third of the lanes idle with one row per wavefront).
struct FlatWalk {
  int r, c, dr, dc, cols;
  __device__ FlatWalk(int cols_) : cols(cols_) {
    const int nthr = blockDim.x * blockDim.y;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    r = tid / cols;
    c = tid - r * cols;
    dr = nthr / cols;
    dc = nthr - dr * cols;
  }
  __device__ void next() {
    c += dc;
    r += dr;
    if (c >= cols) {
      c -= cols;
      r++;
    }
  }
};

// Loads rows [0, R] (row R: the north faces in scz only) x columns [0, C)
// of the tile, the points dealt out by FlatWalk, LOAD_BATCH points per
// thread at a time: every global load of a batch is issued, at clamped
// in-range addresses, before the batch's LDS stores; values outside the
// grid are stored as zero.  With UP (k_up) the prolongation is folded in:
// su = u + (P e) at every point the bilinear prolongation reaches
// (prolongate_kernel, odeSolver.cpp:1017-1043; the same single addition per
// point as a separate pass over su would make), so the coarse correction's
// loads are in flight with the tile's.
template <typename T, bool HAS_F, bool UP, bool FULL_TOP, int LOAD_BATCH>
__device__ inline void tile_load_batches(Tile<T> &t, const T *__restrict__ u,
                                         const T *__restrict__ f,
                                         const T *__restrict__ cx,
                                         const T *__restrict__ cz, int N, int K,
                                         int P, bool zero_u,
                                         const T *__restrict__ e, int Nc, int Kc,
                                         int Pc) {
  const int rows = t.R + 1;
  FlatWalk w(t.C);
  while (w.r < rows) {
    int lr[LOAD_BATCH], lc[LOAD_BATCH];
#pragma unroll
    for (int i = 0; i < LOAD_BATCH; i++) {
      lr[i] = w.r;
      lc[i] = w.c;
      w.next();
    }
    T vu[LOAD_BATCH], vf[LOAD_BATCH], vx[LOAD_BATCH], vz[LOAD_BATCH];
    T e00[LOAD_BATCH], e01[LOAD_BATCH], e10[LOAD_BATCH], e11[LOAD_BATCH];
#pragma unroll
    for (int i = 0; i < LOAD_BATCH; i++) {
      const int j = t.j0 + lr[i], k = t.k0 + lc[i];
      const int jc = min(max(j, 0), N - 1), jz = min(max(j, 0), N);
      const int kc = min(max(k, 0), K - 1);
      const size_t g = (size_t)jc * P + kc;
      vu[i] = u[g];
      if (HAS_F) vf[i] = f[g];
      vx[i] = cx[g];
      vz[i] = cz[(size_t)jz * P + kc];
      if (UP) {
        // coarse point (cj, ck) and its neighbours; clamped like the fine
        // point, so every address is inside the coarse grid
        const int cj = min(jc >> 1, Nc - 1), ck = min(kc >> 1, Kc - 1);
        const int cj1 = min(cj + (jc & 1), Nc - 1), ck1 = min(ck + (kc & 1), Kc - 1);
        e00[i] = e[(size_t)cj * Pc + ck];
        e01[i] = e[(size_t)cj * Pc + ck1];
        e10[i] = e[(size_t)cj1 * Pc + ck];
        e11[i] = e[(size_t)cj1 * Pc + ck1];
      }
    }
#pragma unroll
    for (int i = 0; i < LOAD_BATCH; i++) {
      if (lr[i] >= rows) continue;
      const int j = t.j0 + lr[i], k = t.k0 + lc[i];
      const bool kin = (k >= 0) && (k < K);
      const bool in = kin && (j >= 0) && (j < N) && lr[i] < t.R;
      const bool inz = kin && (j >= 0) && (j <= N);
      const int s = lr[i] * t.LS + lc[i];
      if (lr[i] < t.R) {
        T v = (in && !zero_u) ? vu[i] : T(0);
        if (UP && in) {
          const int cj = j >> 1, dj = j & 1, ck = k >> 1, dk = k & 1;
          const bool pr = (cj < Nc - 1 || (FULL_TOP && cj == Nc - 1 && dj == 0)) &&
                          ck < Kc - 1;
          if (pr) {
            T add;
            if (!dj && !dk)
              add = e00[i];
            else if (!dj)
              add = T(0.5) * (e00[i] + e01[i]);
            else if (!dk)
              add = T(0.5) * (e00[i] + e10[i]);
            else
              add = T(0.25) * (e00[i] + e01[i] + e10[i] + e11[i]);
            v += add;
          }
        }
        t.su[s] = v;
        if (HAS_F) t.sf[s] = in ? vf[i] : T(0);
        t.scx[s] = in ? vx[i] : T(0);
      }
      t.scz[s] = inz ? vz[i] : T(0);
    }
  }
}

// The batch size follows the points per thread (uniform over the
// workgroup): small tiles do not issue loads for empty batch slots.
template <typename T, bool HAS_F, bool UP, bool FULL_TOP>
__device__ inline void tile_load_impl(Tile<T> &t, const T *__restrict__ u,
                                      const T *__restrict__ f,
                                      const T *__restrict__ cx,
                                      const T *__restrict__ cz, int N, int K,
                                      int P, bool zero_u, const T *__restrict__ e,
                                      int Nc, int Kc, int Pc) {
  const int nthr = blockDim.x * blockDim.y;
  const int per = ((t.R + 1) * t.C + nthr - 1) / nthr;
  if (per >= 4)
    tile_load_batches<T, HAS_F, UP, FULL_TOP, 4>(t, u, f, cx, cz, N, K, P, zero_u, e,
                                                 Nc, Kc, Pc);
  else if (per >= 2)
    tile_load_batches<T, HAS_F, UP, FULL_TOP, 2>(t, u, f, cx, cz, N, K, P, zero_u, e,
                                                 Nc, Kc, Pc);
  else
    tile_load_batches<T, HAS_F, UP, FULL_TOP, 1>(t, u, f, cx, cz, N, K, P, zero_u, e,
                                                 Nc, Kc, Pc);
}

template <typename T, bool HAS_F>
__device__ inline void tile_load(Tile<T> &t, const T *__restrict__ u,
                                 const T *__restrict__ f,
                                 const T *__restrict__ cx,
                                 const T *__restrict__ cz, int N, int K, int P,
                                 bool zero_u) {
  tile_load_impl<T, HAS_F, false, false>(t, u, f, cx, cz, N, K, P, zero_u, nullptr,
                                         0, 0, 0);
}

template <typename T, bool HAS_F, bool FULL_TOP>
__device__ inline void tile_load_up(Tile<T> &t, const T *__restrict__ u,
                                    const T *__restrict__ f,
                                    const T *__restrict__ cx,
                                    const T *__restrict__ cz, int N, int K, int P,
                                    const T *__restrict__ e, int Nc, int Kc,
                                    int Pc) {
  tile_load_impl<T, HAS_F, true, FULL_TOP>(t, u, f, cx, cz, N, K, P, false, e, Nc,
                                           Kc, Pc);
}

// Half-sweep `colour` over the local window [h, R-h) x [h, C-h): the half
// points of one colour in each row, all rows, dealt out by FlatWalk.  The
// points of one colour are independent of each other, so the order in
// which they are updated does not change the result.
template <typename T, bool HAS_F, bool EXACT>
__device__ inline void tile_half_sweep(Tile<T> &t, int h, int colour, int N,
                                       int K, T DH2, T relax) {
  const int cols = t.C - 2 * h;
  const int half = (cols + 1) / 2;
  const int rows = t.R - 2 * h;
  for (FlatWalk w(half); w.r < rows; w.next()) {
    const int lr = h + w.r;
    const int j = t.j0 + lr;
    const int delta = (colour + j + t.k0 + h) & 1;
    const int lc = h + 2 * w.c + delta;
    const int k = t.k0 + lc;
    if (j < 0 || j >= N || lc >= t.C - h || k < 1 || k > K - 2) continue;
    const int lS = (j == 0) ? lr + 1 : lr - 1;
    const int lN = (j == N - 1) ? lr - 1 : lr + 1;
    const int s = lr * t.LS + lc;
    const T m1 = t.scx[s], m2 = t.scx[s + 1], m3 = t.scz[s], m4 = t.scz[s + t.LS];
    const T fv = HAS_F ? DH2 * t.sf[s] : T(0);
    t.su[s] = rbgs_update<T, EXACT>(m1, m2, m3, m4, t.su[s - 1], t.su[s + 1],
                                    t.su[lS * t.LS + lc], t.su[lN * t.LS + lc],
                                    t.su[s], fv, HAS_F, relax);
  }
}

// The sweeps with each thread holding fixed 2x2 quads of the tile (local
// rows r, r+1 and columns c, c+1, r and c even) in registers: their values,
// coefficients and right-hand sides.  The tile's origin (j0, k0) has even
// parity (even tile sizes, halo on both sides), so in every quad (r, c) and
// (r+1, c+1) are colour 0 and the other two colour 1, for every thread alike.
// A point's neighbours are two of its own quad (registers) and two outside
// it (LDS): two LDS reads per update instead of nine.  Every update is also
// stored to the tile, which stays current for the other threads and for
// what follows the sweeps.  Values, coefficients and expressions are those
// of tile_half_sweep, so the results are the same bits.
template <typename T, bool HAS_F>
struct Quads1 {
  T u00, u01, u10, u11;
  T x0, x1, x2;  // scx row r, columns c .. c+2
  T y0, y1, y2;  // scx row r+1
  T z0, z1;      // scz row r (south faces of row r)
  T w0, w1;      // scz row r+1
  T v0, v1;      // scz row r+2
  T f00, f01, f10, f11;  // DH^2 f (HAS_F)
  // per-point relaxation factor relax / (m1 + m2 + m3 + m4) of the exact
  // update, or the coefficient sum m1 + m2 + m3 + m4 of the plain one, when
  // the quads keep it (PRE)
  T a00, a01, a10, a11;
  int r, c;
  bool valid;
};

// The update of rbgs_update with the per-point quantity a formed once:
// relax / (m1 + m2 + m3 + m4) (EXACT) or m1 + m2 + m3 + m4 (plain), the
// same operations in the same order, so the same bits.
template <typename T, bool EXACT>
__device__ inline T rbgs_update_pre(T a, T m1, T m2, T m3, T m4, T uW, T uE, T uS,
                                    T uN, T uC, T fDH2, bool has_f, T relax) {
  T num = m1 * uW + m2 * uE + m3 * uS + m4 * uN;
  if (has_f) num = num - fDH2;
  if (EXACT) return a * num + (T(1) - relax) * uC;
  const T unew = num / a;
  return (relax == T(1)) ? unew : uC + relax * (unew - uC);
}

template <typename T, bool EXACT>
__device__ inline T rbgs_factor(T m1, T m2, T m3, T m4, T relax) {
  return EXACT ? relax / (m1 + m2 + m3 + m4) : m1 + m2 + m3 + m4;
}

// Point P of quad q in the half-sweep of colour COL: P = 0 the upper point
// of the colour (row r), P = 1 the lower (row r+1).
template <typename T, bool HAS_F, bool EXACT, int COL, int P>
__device__ inline void quad_point_read(const Tile<T> &t, const Quads1<T, HAS_F> &q,
                                       int h, int N, int K, bool &ok, int &at,
                                       T &hout, T &vout) {
  const bool top = P == 0;
  const bool left = (COL == 0) == (P == 0);
  const int lr = q.r + P;
  const int lc = q.c + (left ? 0 : 1);
  const int j = t.j0 + lr, k = t.k0 + lc;
  ok = q.valid && lr >= h && lr < t.R - h && lc >= h && lc < t.C - h && j >= 0 &&
       j < N && k >= 1 && k <= K - 2;
  at = lr * t.LS + lc;
  const int ih = ok ? at + (left ? -1 : 1) : t.LS + 1;
  const int iv = ok ? at + (top ? -t.LS : t.LS) : t.LS + 1;
  hout = t.su[ih];
  vout = t.su[iv];
}

template <typename T, bool HAS_F, bool EXACT, int COL, int P, bool PRE>
__device__ inline T quad_point_update(const Tile<T> &t, const Quads1<T, HAS_F> &q,
                                      int N, T hout, T vout, T relax) {
  const int j = t.j0 + q.r + P;
  T m1, m2, m3, m4, uW, uE, uC, vin, fv, a;
  if (COL == 0 && P == 0) {  // (r, c)
    m1 = q.x0; m2 = q.x1; m3 = q.z0; m4 = q.w0;
    uW = hout; uE = q.u01; uC = q.u00; vin = q.u10; fv = q.f00; a = q.a00;
  } else if (COL == 0) {     // (r+1, c+1)
    m1 = q.y1; m2 = q.y2; m3 = q.w1; m4 = q.v1;
    uW = q.u10; uE = hout; uC = q.u11; vin = q.u01; fv = q.f11; a = q.a11;
  } else if (P == 0) {       // (r, c+1)
    m1 = q.x1; m2 = q.x2; m3 = q.z1; m4 = q.w1;
    uW = q.u00; uE = hout; uC = q.u01; vin = q.u11; fv = q.f01; a = q.a01;
  } else {                   // (r+1, c)
    m1 = q.y0; m2 = q.y1; m3 = q.w0; m4 = q.v0;
    uW = hout; uE = q.u11; uC = q.u10; vin = q.u00; fv = q.f10; a = q.a10;
  }
  // south/north neighbours; rows 0 and N-1 of the grid reflect (the
  // neighbour across the boundary is the one on the other side)
  T uS, uN;
  if (P == 0) {
    uS = (j == 0) ? vin : vout;
    uN = (j == N - 1) ? vout : vin;
  } else {
    uS = (j == 0) ? vout : vin;
    uN = (j == N - 1) ? vin : vout;
  }
  if (PRE)
    return rbgs_update_pre<T, EXACT>(a, m1, m2, m3, m4, uW, uE, uS, uN, uC,
                                     HAS_F ? fv : T(0), HAS_F, relax);
  return rbgs_update<T, EXACT>(m1, m2, m3, m4, uW, uE, uS, uN, uC,
                               HAS_F ? fv : T(0), HAS_F, relax);
}

// One quad at a time: its two points' reads, updates, then stores.
template <typename T, bool HAS_F, bool EXACT, int COL, int S, bool PRE>
__device__ inline void quad_half_sweep(Tile<T> &t, Quads1<T, HAS_F> (&q)[S], int h,
                                       int N, int K, T relax) {
#pragma unroll
  for (int i = 0; i < S; i++) {
    bool ok0, ok1;
    int at0, at1;
    T h0, v0, h1, v1;
    quad_point_read<T, HAS_F, EXACT, COL, 0>(t, q[i], h, N, K, ok0, at0, h0, v0);
    quad_point_read<T, HAS_F, EXACT, COL, 1>(t, q[i], h, N, K, ok1, at1, h1, v1);
    const T n0 =
        quad_point_update<T, HAS_F, EXACT, COL, 0, PRE>(t, q[i], N, h0, v0, relax);
    const T n1 =
        quad_point_update<T, HAS_F, EXACT, COL, 1, PRE>(t, q[i], N, h1, v1, relax);
    if (ok0) {
      t.su[at0] = n0;
      if (COL == 0) q[i].u00 = n0; else q[i].u01 = n0;
    }
    if (ok1) {
      t.su[at1] = n1;
      if (COL == 0) q[i].u11 = n1; else q[i].u10 = n1;
    }
  }
}

// Whether the sweeps run on register quads, and with how many per thread
// (0: the point walk): two or more sweeps (loading the registers costs
// about one half-sweep of LDS reads), up to two quads per thread (one with
// a right-hand side in FP64).
template <typename T, bool HAS_F>
__device__ inline int quads_per_thread(const Tile<T> &t, int nu) {
  const int nthr = blockDim.x * blockDim.y;
  const int nq = (t.R >> 1) * (t.C >> 1);
  if (((t.R | t.C) & 1) != 0 || nu < 2) return 0;
  if (nq <= nthr) return 1;
  if (nq <= 2 * nthr && (!HAS_F || sizeof(T) == 4)) return 2;
  return 0;
}

template <typename T, bool HAS_F, bool EXACT, int S>
__device__ inline void tile_sweeps_quads(Tile<T> &t, int nu, int N, int K, T DH2,
                                         T relax) {
  // the per-point factor in registers too where they allow it: one quad
  // per thread, or FP32 values (two FP64 quads already fill the registers)
  constexpr bool PRE = S == 1 || sizeof(T) == 4;
  const int nthr = blockDim.x * blockDim.y;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  const int qc = t.C >> 1, nq = (t.R >> 1) * qc;
  const int LS = t.LS;
  Quads1<T, HAS_F> q[S];
#pragma unroll
  for (int i = 0; i < S; i++) {
    const int id = tid + i * nthr;
    q[i].valid = id < nq;
    const int a = q[i].valid ? id / qc : 0;
    const int b = q[i].valid ? id - a * qc : 0;
    q[i].r = 2 * a;
    q[i].c = 2 * b;
    // rows r .. r+2 of scz and column c+2 of scx stay inside the tile's
    // LDS (scz has R+1 rows; column C of a row is the next row's first)
    const int s = q[i].r * LS + q[i].c;
    q[i].u00 = t.su[s]; q[i].u01 = t.su[s + 1];
    q[i].u10 = t.su[s + LS]; q[i].u11 = t.su[s + LS + 1];
    q[i].x0 = t.scx[s]; q[i].x1 = t.scx[s + 1]; q[i].x2 = t.scx[s + 2];
    q[i].y0 = t.scx[s + LS]; q[i].y1 = t.scx[s + LS + 1]; q[i].y2 = t.scx[s + LS + 2];
    q[i].z0 = t.scz[s]; q[i].z1 = t.scz[s + 1];
    q[i].w0 = t.scz[s + LS]; q[i].w1 = t.scz[s + LS + 1];
    q[i].v0 = t.scz[s + 2 * LS]; q[i].v1 = t.scz[s + 2 * LS + 1];
    if (HAS_F) {
      q[i].f00 = DH2 * t.sf[s]; q[i].f01 = DH2 * t.sf[s + 1];
      q[i].f10 = DH2 * t.sf[s + LS]; q[i].f11 = DH2 * t.sf[s + LS + 1];
    } else {
      q[i].f00 = q[i].f01 = q[i].f10 = q[i].f11 = T(0);
    }
    if (PRE) {
      q[i].a00 = rbgs_factor<T, EXACT>(q[i].x0, q[i].x1, q[i].z0, q[i].w0, relax);
      q[i].a01 = rbgs_factor<T, EXACT>(q[i].x1, q[i].x2, q[i].z1, q[i].w1, relax);
      q[i].a10 = rbgs_factor<T, EXACT>(q[i].y0, q[i].y1, q[i].w0, q[i].v0, relax);
      q[i].a11 = rbgs_factor<T, EXACT>(q[i].y1, q[i].y2, q[i].w1, q[i].v1, relax);
    } else {
      q[i].a00 = q[i].a01 = q[i].a10 = q[i].a11 = T(0);
    }
  }
  for (int s = 0; s < 2 * nu; s += 2) {
    quad_half_sweep<T, HAS_F, EXACT, 0, S, PRE>(t, q, s + 1, N, K, relax);
    __syncthreads();
    quad_half_sweep<T, HAS_F, EXACT, 1, S, PRE>(t, q, s + 2, N, K, relax);
    __syncthreads();
  }
}

template <typename T, bool HAS_F, bool EXACT>
__device__ inline void tile_sweeps(Tile<T> &t, int nu, int N, int K, T DH2,
                                   T relax) {
  const int qs = quads_per_thread<T, HAS_F>(t, nu);
  if (qs == 1) {
    tile_sweeps_quads<T, HAS_F, EXACT, 1>(t, nu, N, K, DH2, relax);
    return;
  }
  if (qs == 2) {
    tile_sweeps_quads<T, HAS_F, EXACT, 2>(t, nu, N, K, DH2, relax);
    return;
  }
  for (int s = 0; s < 2 * nu; s++) {
    tile_half_sweep<T, HAS_F, EXACT>(t, s + 1, s & 1, N, K, DH2, relax);
    __syncthreads();
  }
}

// residual of local point (lr, lc) (global row j), same expression as
// k_residual.
template <typename T, bool HAS_F>
__device__ inline T tile_residual(const Tile<T> &t, int lr, int lc, int j,
                                  int N, T DH2) {
  const int lS = (j == 0) ? lr + 1 : lr - 1;
  const int lN = (j == N - 1) ? lr - 1 : lr + 1;
  const int s = lr * t.LS + lc;
  const T m1 = t.scx[s], m2 = t.scx[s + 1], m3 = t.scz[s], m4 = t.scz[s + t.LS];
  const T fv = HAS_F ? t.sf[s] : T(0);
  return fv - (m1 * t.su[s - 1] + m2 * t.su[s + 1] + m3 * t.su[lS * t.LS + lc] +
               m4 * t.su[lN * t.LS + lc] - (m1 + m2 + m3 + m4) * t.su[s]) /
                  DH2;
}

// Tile handled by this workgroup.  Workgroups are dispatched to the 8 XCDs
// in turn (linear id mod 8), and each XCD has its own L2; the tiles are
// dealt out in 8 contiguous bands of row-major tile order, one per XCD, so
// neighbouring tiles (their halos overlap) and the tiles of consecutive
// levels over the same part of the grid run on the same XCD.  XCD x holds
// q + (x < r) workgroups for n = 8 q + r tiles.
__device__ inline int tile_id() {
  const int n = gridDim.x * gridDim.y;
  const int id = blockIdx.y * gridDim.x + blockIdx.x;
  const int x = id & 7, slot = id >> 3;
  const int q = n >> 3, r = n & 7;
  return x * q + min(x, r) + slot;
}

template <typename T>
__device__ inline Tile<T> make_tile(unsigned char *smem, int TX, int TY, int H,
                                    bool has_f) {
  Tile<T> t;
  t.R = TY + 2 * H;
  t.C = TX + 2 * H;
  t.LS = t.C;
  const int tile = tile_id();
  t.j0 = (tile / (int)gridDim.x) * TY - H;
  t.k0 = (tile % (int)gridDim.x) * TX - H;
  T *p = reinterpret_cast<T *>(smem);
  const int n = t.R * t.LS;
  t.su = p;
  t.scx = p + n;
  t.scz = p + 2 * n;
  t.sf = has_f ? p + 3 * n + t.LS : nullptr;
  return t;
}

// Without a right-hand side (level 0 of the FP64 cycle) k_down keeps the
// residuals of its (TY+2) x (TX+2) window in registers, up to RES_REGS per
// thread, and then writes them over scx, which the restriction no longer
// needs: the tile then takes three arrays of LDS instead of four, which
// lets two workgroups share a CU.  Larger windows use a scratch array.
constexpr int RES_REGS = 8;
__host__ __device__ inline bool res_in_regs(int TX, int TY, int threads) {
  return (TX + 2) * (TY + 2) <= RES_REGS * threads;
}

// Start of a plain FP64 solve folded into its first V-cycle (no separate
// start kernels): the level-0 down kernel (START_PARTIALS) forms, per tile
// and before smoothing, the residual-norm partials of the start and of the
// boundary data (the residual with the interior at zero), the expressions of
// k_start_partials; the level-1 down kernel (START_TEST) sums them in a
// fixed order (every workgroup the same bits), applies the start test of
// k_start_gate, returns at once if the start already meets the threshold,
// and its workgroup (0, 0) publishes the start as k_start_gate does.  Both
// ignore the gate flag, which the previous solve left set.
enum { DOWN_PLAIN = 0, START_PARTIALS = 1, START_TEST = 2 };
// partials in c->partial: start norm at kStartOff + tile, boundary data at
// kStartOff + kStartCap + tile (the level-0 up kernel's norm partials stay
// below kStartOff)
constexpr int kStartOff = 8192, kStartCap = 4096;
static_assert(kStartOff + 2 * kStartCap <= kPartialCap, "start partials");

struct StartArgs {
  double *part;    // c->partial + kStartOff
  int nb;          // tiles of the level-0 down kernel (partials per kind)
  double *scal, *mb;
  int *gate;
  int rel, cap;
  double tol;
  // launched ahead (spec: the gate word it waits on, else 0): START_PARTIALS
  // goes on only if gate[spec] == 1 (2: the solve before it converged; 7:
  // k_combine accepted the attempt) and records that decision in gate[6];
  // START_TEST abandons the start when gate[6] is 0 (gate[0] = 1 stops the
  // batch behind it, the mailbox step carries outcome 4)
  int spec;
  // DOWN_PLAIN at level 0: the stop test of the previous V-cycle (that of
  // k_norm_gate, on its nb <= 256 norm partials at part) runs here
  int fold;
};

// Pre-smoothing + residual + restriction (levels above the tail).
template <typename T, bool HAS_F, bool EXACT, int SMODE>
__global__ __launch_bounds__(TILE_THREADS) void k_down(
    const T *__restrict__ uin, T *__restrict__ uout, const T *__restrict__ f,
    const T *__restrict__ cx, const T *__restrict__ cz, int N, int K, int P,
    T DH2, T relax, int nu, int zero_u, T *__restrict__ r2, int Nc, int Kc,
    int Pc, int TX, int TY, const int *__restrict__ gate, StartArgs sa) {
  const bool wg0 = blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 &&
                   threadIdx.y == 0;
  const bool folding = SMODE == DOWN_PLAIN && !HAS_F && sa.fold;
  if (SMODE == DOWN_PLAIN && gate && gate[0]) {  // converged: skip
    // a folded stop test still takes its step of the mailbox sequence
    if (folding && wg0) sa.gate[3] = sa.gate[3] + 1;
    return;
  }
  if (SMODE == START_PARTIALS && sa.spec) {
    const int ok = sa.gate[sa.spec] == 1 ? 1 : 0;
    if (wg0) sa.gate[6] = ok;
    if (!ok) return;
  }
  if (SMODE == START_TEST && sa.spec && sa.gate[6] == 0) {
    if (wg0) {
      sa.gate[0] = 1;
      const int seq = sa.gate[3] + 1;
      sa.gate[3] = seq;
      sa.mb[MB_CODE] = 4.0;
      __threadfence_system();
      sa.mb[MB_SEQ] = (double)seq;
    }
    return;
  }
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double sdec[16 * 2];
  if (SMODE == START_TEST && threadIdx.y == 0) {
    // wavefront 0: the start norms from the level-0 tiles' partials
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < sa.nb; i += 64) {
      a += sa.part[i];
      b += sa.part[kStartCap + i];
    }
    a = wave_sum(a);
    b = wave_sum(b);
    if (threadIdx.x == 0) {
      const double du = sqrt(a);
      const double r0 = sa.rel ? sqrt(b) : du;
      const double thr = sa.rel ? sa.tol * r0 : sa.tol;
      int code = 0;
      if (!(du > thr))
        code = 1;
      else if (sa.cap < 1)
        code = 3;
      sdec[0] = (double)code;
      if (blockIdx.x == 0 && blockIdx.y == 0) {
        sa.scal[SC_NORM] = du;
        sa.scal[SC_NORM0] = r0;
        sa.gate[0] = code ? 1 : 0;
        sa.gate[1] = 0;
        sa.gate[2] = code;
        sa.gate[4] = 0;
        sa.gate[5] = 0;
        sa.mb[SC_NORM] = du;
        sa.mb[SC_NORM0] = r0;
        sa.mb[MB_START] = du;
        sa.mb[MB_USED] = 0.0;
        sa.mb[MB_COUNT] = 0.0;
        sa.mb[MB_CODE] = (double)code;
        const int seq = sa.gate[3] + 1;
        sa.gate[3] = seq;
        __threadfence_system();
        sa.mb[MB_SEQ] = (double)seq;
      }
    }
  }
  const int H = 2 * nu + 2;
  Tile<T> t = make_tile<T>(smem, TX, TY, H, HAS_F);
  // folded stop test: each of k_norm_gate's 256 threads (the first four
  // wavefronts here) loads its partial before the tile load and sums after
  // it, in k_norm_gate's order, so every workgroup takes the same decision
  const int ltid = threadIdx.y * blockDim.x + threadIdx.x;
  double fold_p = 0.0;
  if (folding && ltid < sa.nb) fold_p = sa.part[ltid];
  tile_load<T, HAS_F>(t, uin, f, cx, cz, N, K, P, zero_u != 0);
  if (folding && ltid < 256) {
    const double w = wave_sum(0.0 + fold_p);
    if ((ltid & 63) == 0) sdec[ltid >> 6] = w;
  }
  __syncthreads();
  if (SMODE == START_TEST && sdec[0] != 0.0) return;  // no V-cycle needed
  if (folding) {
    const double v = sqrt(((sdec[0] + sdec[1]) + sdec[2]) + sdec[3]);
    const double thr = sa.rel ? sa.tol * sa.scal[SC_NORM0] : sa.tol;
    int code = 0;
    if (isnan(v) || isinf(v))
      code = 2;
    else if (!(v > thr))
      code = 1;
    if (wg0) {
      // what k_norm_gate records (the V-cycle cap is left to it: the host
      // folds no test that could reach the cap)
      const int cnt = sa.gate[1] + 1;
      sa.gate[1] = cnt;
      sa.gate[2] = code;
      sa.gate[3] = sa.gate[3] + 1;
      sa.scal[SC_NORM] = v;
      if (code) {
        sa.gate[0] = 1;
        sa.mb[SC_NORM] = v;
        sa.mb[MB_COUNT] = (double)cnt;
        sa.mb[MB_CODE] = (double)code;
        __threadfence_system();
      }
    }
    if (code) return;
  }
  if (SMODE == START_PARTIALS) {
    // the start as loaded: residual norm and boundary-data norm partials
    double v2[2] = {0.0, 0.0};
    for (FlatWalk w(TX); w.r < TY; w.next()) {
      const int lr = H + w.r, lc = H + w.c;
      const int j = t.j0 + lr, k = t.k0 + lc;
      if (j >= N || k < 1 || k > K - 2) continue;
      const double r = (double)tile_residual<T, false>(t, lr, lc, j, N, DH2);
      v2[0] += r * r;
      if (k == 1 || k == K - 2) {
        const int s = lr * t.LS + lc;
        const T uW = (k - 1 == 0) ? t.su[s - 1] : T(0);
        const T uE = (k + 1 == K - 1) ? t.su[s + 1] : T(0);
        const T rb = T(0) - (t.scx[s] * uW + t.scx[s + 1] * uE) / DH2;
        v2[1] += (double)(rb * rb);
      }
    }
    block_sums<2>(v2, sdec);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
      const int bid = tile_id();
      sa.part[bid] = v2[0];
      sa.part[kStartCap + bid] = v2[1];
    }
  }
  tile_sweeps<T, HAS_F, EXACT>(t, nu, N, K, DH2, relax);
  // residual on the interior plus one ring; it replaces f (read at the same
  // point only), or goes over scx (after a barrier, from registers), or to
  // the scratch block after scz
  const int h = H - 1;
  const int rrows = t.R - 2 * h, rcols = t.C - 2 * h;
  T *sr;
  if (!HAS_F && res_in_regs(TX, TY, blockDim.x * blockDim.y)) {
    T rv[RES_REGS];
    {
      FlatWalk w(rcols);
#pragma unroll
      for (int i = 0; i < RES_REGS; i++) {
        rv[i] = T(0);
        if (w.r < rrows) {
          const int lr = h + w.r, lc = h + w.c;
          const int j = t.j0 + lr;
          const int k = t.k0 + lc;
          if (j >= 0 && j < N && k >= 1 && k <= K - 2)
            rv[i] = tile_residual<T, HAS_F>(t, lr, lc, j, N, DH2);
        }
        w.next();
      }
    }
    __syncthreads();
    sr = t.scx;
    {
      FlatWalk w(rcols);
#pragma unroll
      for (int i = 0; i < RES_REGS; i++) {
        if (w.r < rrows) sr[(h + w.r) * t.LS + h + w.c] = rv[i];
        w.next();
      }
    }
  } else {
    sr = HAS_F ? t.sf : t.scz + (t.R + 1) * t.LS;
    for (FlatWalk w(rcols); w.r < rrows; w.next()) {
      const int lr = h + w.r, lc = h + w.c;
      const int j = t.j0 + lr;
      const int k = t.k0 + lc;
      T r = T(0);
      if (j >= 0 && j < N && k >= 1 && k <= K - 2)
        r = tile_residual<T, HAS_F>(t, lr, lc, j, N, DH2);
      sr[lr * t.LS + lc] = r;
    }
  }
  __syncthreads();
  // restriction for the coarse points inside the tile interior
  const int jc0 = (t.j0 + H) / 2, kc0 = (t.k0 + H) / 2;
  for (FlatWalk w(TX / 2); w.r < TY / 2; w.next()) {
    const int jc = jc0 + w.r;
    const int kc = kc0 + w.c;
    if (jc >= Nc || kc < 1 || kc > Kc - 2) continue;
    const int jS = (jc == 0) ? 1 : 2 * jc - 1;
    const int jN = (jc == Nc - 1) ? 2 * jc - 1 : 2 * jc + 1;
    const T *rc = sr + (2 * jc - t.j0) * t.LS - t.k0;
    const T *rs = sr + (jS - t.j0) * t.LS - t.k0;
    const T *rn = sr + (jN - t.j0) * t.LS - t.k0;
    const int kf = 2 * kc;
    const T v = T(0.0625) *
                (T(4.0) * rc[kf] + T(2.0) * rc[kf - 1] + T(2.0) * rc[kf + 1] +
                 T(2.0) * rs[kf] + T(2.0) * rn[kf] + rs[kf - 1] + rs[kf + 1] +
                 rn[kf - 1] + rn[kf + 1]);
    r2[(size_t)jc * Pc + kc] = v;
  }
  // smoothed interior
  for (FlatWalk w(TX); w.r < TY; w.next()) {
    const int lr = H + w.r, lc = H + w.c;
    const int j = t.j0 + lr;
    const int k = t.k0 + lc;
    if (j < N && k < K) uout[(size_t)j * P + k] = t.su[lr * t.LS + lc];
  }
}

// Level-0 finishing modes of k_up.
enum { UP_STORE = 0, UP_NORM = 1, UP_ADD64 = 2 };

// Prolongation + correction + post-smoothing.
template <typename T, bool HAS_F, bool EXACT, bool FULL_TOP, int MODE>
__global__ __launch_bounds__(TILE_THREADS) void k_up(
    const T *__restrict__ uin, T *__restrict__ uout, const T *__restrict__ f,
    const T *__restrict__ cx, const T *__restrict__ cz, int N, int K, int P,
    T DH2, T relax, int nu, const T *__restrict__ e, int Nc, int Kc, int Pc,
    int TX, int TY, double *__restrict__ partial, double *__restrict__ u64,
    const int *__restrict__ gate) {
  if (gate && gate[0]) return;  // the solve has converged: skip this cycle
  extern __shared__ __align__(16) unsigned char smem[];
  const int H = 2 * nu + (MODE == UP_NORM ? 1 : 0);
  Tile<T> t = make_tile<T>(smem, TX, TY, H, HAS_F);
  // u + P e (prolongate_kernel, odeSolver.cpp:1017-1043), loaded together
  tile_load_up<T, HAS_F, FULL_TOP>(t, uin, f, cx, cz, N, K, P, e, Nc, Kc, Pc);
  __syncthreads();
  tile_sweeps<T, HAS_F, EXACT>(t, nu, N, K, DH2, relax);
  if (MODE == UP_NORM) {
    // the residual-norm partial of the tile's interior first: the block
    // sum's barrier then has no global stores of this kernel to wait for
    double acc = 0.0;
    for (FlatWalk w(TX); w.r < TY; w.next()) {
      const int lr = H + w.r, lc = H + w.c;
      const int j = t.j0 + lr;
      const int k = t.k0 + lc;
      if (j < N && k >= 1 && k <= K - 2) {
        const double r = (double)tile_residual<T, HAS_F>(t, lr, lc, j, N, DH2);
        acc += r * r;
      }
    }
    __shared__ double red[16];
    const double s = block_sum1(acc, red);
    if (threadIdx.x == 0 && threadIdx.y == 0)
      partial[tile_id()] = s;
  }
  for (FlatWalk w(TX); w.r < TY; w.next()) {
    const int lr = H + w.r, lc = H + w.c;
    const int j = t.j0 + lr;
    const int k = t.k0 + lc;
    if (j < N && k < K) {
      const size_t g = (size_t)j * P + k;
      if (MODE == UP_ADD64) {
        if (k >= 1 && k <= K - 2) u64[g] += (double)t.su[lr * t.LS + lc];
      } else {
        uout[g] = t.su[lr * t.LS + lc];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Coarse tail: one workgroup, every level (values and coefficients) in LDS.
// ---------------------------------------------------------------------------

struct TailLevel {
  const void *cx, *cz;  // global coefficients of the level
  int N, K, P;          // dims and global pitch
  double DH2;
};

struct TailArgs {
  TailLevel lv[6];
  int nlev;             // levels in the tail (first = entry level)
  const void *f_in;     // restricted residual of the entry level (global)
  void *e_out;          // correction of the entry level (global)
  int nu_pre, nu_post, nu_coarse;
  double relax;
  const void *minv;     // direct coarsest solve (column-major FP32 inverse) or null
  int cn;               // coarsest unknowns for minv
  const int *gate;      // skip flag of a batched solve, or null
};

// compact arrays of a tail level, stride K: u, f, cx and cz (N+1 rows)
template <typename T>
struct TailBufs {
  T *u, *f, *cx, *cz;
};

__host__ __device__ inline size_t tail_level_words(int N, int K) {
  return 4 * (size_t)N * K + K;
}

// LDS arrays of tail level l, recomputed from the level dimensions (uniform
// values) at each use: a per-level table indexed by a loop variable would
// live in per-lane scratch memory.
template <typename T>
__device__ inline TailBufs<T> tail_bufs(T *base, const TailArgs &a, int l) {
  size_t off = 0;
  for (int i = 0; i < l; i++) off += tail_level_words(a.lv[i].N, a.lv[i].K);
  const size_t n = (size_t)a.lv[l].N * a.lv[l].K;
  T *b = base + off;
  return TailBufs<T>{b, b + n, b + 2 * n, b + 3 * n};
}

template <typename T>
__device__ inline T *tail_end(T *base, const TailArgs &a) {
  size_t off = 0;
  for (int i = 0; i < a.nlev; i++) off += tail_level_words(a.lv[i].N, a.lv[i].K);
  return base + off;
}

// Sweeps on one tail level.  With one_wave, the first wavefront alone runs
// them (small coarsest grids), ordering its own LDS traffic between
// half-sweeps with wavefront fences instead of workgroup barriers.
// One point of colour COL of a thread that holds its coefficients, DH^2 f
// and relaxation factor in registers.
template <typename T>
struct TailPoint {
  T m1, m2, m3, m4, a, fv;
  int c, cS, cN;
  bool ok;
};

template <typename T, bool EXACT>
__device__ inline TailPoint<T> tail_point(const TailBufs<T> &b, int N, int K, T DH2,
                                          T relax, bool act, int j, int q, int col) {
  TailPoint<T> p;
  const int k = 2 * q + ((j + col) & 1);
  p.ok = act && k >= 1 && k <= K - 2;
  const int kk = p.ok ? k : 1;  // a point that stays put reads in-range values
  const int jS = (j == 0) ? 1 : j - 1;
  const int jN = (j == N - 1) ? N - 2 : j + 1;
  p.c = j * K + kk;
  p.cS = jS * K + kk;
  p.cN = jN * K + kk;
  p.m1 = b.cx[p.c];
  p.m2 = b.cx[p.c + 1];
  p.m3 = b.cz[p.c];
  p.m4 = b.cz[p.c + K];
  p.a = rbgs_factor<T, EXACT>(p.m1, p.m2, p.m3, p.m4, relax);
  p.fv = DH2 * b.f[p.c];
  return p;
}

template <typename T, bool EXACT>
__device__ inline void tail_point_update(const TailBufs<T> &b, const TailPoint<T> &p,
                                         T relax) {
  const T v = rbgs_update_pre<T, EXACT>(p.a, p.m1, p.m2, p.m3, p.m4, b.u[p.c - 1],
                                        b.u[p.c + 1], b.u[p.cS], b.u[p.cN], b.u[p.c],
                                        p.fv, true, relax);
  if (p.ok) b.u[p.c] = v;
}

template <typename T, bool EXACT>
__device__ inline void tail_smooth(const TailBufs<T> &b, int N, int K, T DH2,
                                   T relax, int sweeps, bool one_wave) {
  const int half = (K + 1) / 2;
  const int tid = threadIdx.x;
  const int nthr = one_wave ? 64 : (int)blockDim.x;
  if (one_wave && tid >= 64) return;
  if (N * half <= nthr && !one_wave) {
    // one point of each colour per thread: its coefficients, right-hand
    // side and factor read once into registers (f does not change while the
    // level is smoothed), the same update expression
    const bool act = tid < N * half;
    const int j = act ? tid / half : 0, q = act ? tid - j * half : 0;
    const TailPoint<T> p0 = tail_point<T, EXACT>(b, N, K, DH2, relax, act, j, q, 0);
    const TailPoint<T> p1 = tail_point<T, EXACT>(b, N, K, DH2, relax, act, j, q, 1);
    for (int s = 0; s < sweeps; s++) {
      tail_point_update<T, EXACT>(b, p0, relax);
      __syncthreads();
      tail_point_update<T, EXACT>(b, p1, relax);
      __syncthreads();
    }
    return;
  }
  for (int s = 0; s < 2 * sweeps; s++) {
    const int colour = s & 1;
    for (int idx = tid; idx < N * half; idx += nthr) {
      const int j = idx / half, q = idx - j * half;
      const int k = 2 * q + ((j + colour) & 1);
      if (k < 1 || k > K - 2) continue;
      const int jS = (j == 0) ? 1 : j - 1;
      const int jN = (j == N - 1) ? N - 2 : j + 1;
      const int c = j * K + k;
      b.u[c] = rbgs_update<T, EXACT>(b.cx[c], b.cx[c + 1], b.cz[c], b.cz[c + K],
                                     b.u[c - 1], b.u[c + 1], b.u[jS * K + k],
                                     b.u[jN * K + k], b.u[c], DH2 * b.f[c], true,
                                     relax);
    }
    if (one_wave) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
      __syncthreads();
    }
  }
}

template <typename T, bool EXACT, bool FULL_TOP>
__global__ __launch_bounds__(TAIL_THREADS) void k_tail(TailArgs a) {
  if (a.gate && a.gate[0]) return;  // the solve has converged
  extern __shared__ __align__(16) unsigned char smem[];
  T *base = reinterpret_cast<T *>(smem);
  T *Rs = tail_end(base, a);  // residual scratch (entry level size)
  T *Gs = Rs + (a.lv[0].N * a.lv[0].K > a.cn ? a.lv[0].N * a.lv[0].K : a.cn);
  T *Ps = Gs + a.cn;
  // Gs: right-hand side DH^2 f of the direct coarsest solve; Ps: its 1024
  // partial sums
  const int tid = threadIdx.x;
  const T relax = (T)a.relax;
  // Global loads through address-space-1 pointers, so the compiler may keep
  // them in flight across the LDS stores.  The entry level's right-hand side
  // is loaded first, together with the first level's coefficients.
  typedef const __attribute__((address_space(1))) T GT;
  const TailLevel &L0 = a.lv[0];
  T fv[2];
  bool fin[2];
#pragma unroll
  for (int it = 0; it < 2; it++) {
    // clamped indices: every load is issued, unused values are dropped at
    // the LDS store
    const int n0 = L0.N * L0.K;
    const int idx = tid + it * TAIL_THREADS;
    const int ii = idx < n0 ? idx : n0 - 1;
    const int j = ii / L0.K, k = ii - j * L0.K;
    const int kc = k < 1 ? 1 : (k > L0.K - 2 ? L0.K - 2 : k);
    fv[it] = ((GT *)a.f_in)[(size_t)j * L0.P + kc];
    fin[it] = k == kc;
  }
  for (int l = 0; l < a.nlev; l++) {
    const TailLevel &L = a.lv[l];
    const TailBufs<T> b = tail_bufs(base, a, l);
    GT *cx = (GT *)L.cx;
    GT *cz = (GT *)L.cz;
    const int nz = (L.N + 1) * L.K, nx = L.N * L.K;
    T vx[2], vz[2];
#pragma unroll
    for (int it = 0; it < 2; it++) {
      const int idx = tid + it * TAIL_THREADS;
      const int ii = idx < nz ? idx : nz - 1;
      const int j = ii / L.K, k = ii - j * L.K;
      const int jx = j < L.N ? j : L.N - 1;
      vz[it] = cz[(size_t)j * L.P + k];
      vx[it] = cx[(size_t)jx * L.P + k];
    }
#pragma unroll
    for (int it = 0; it < 2; it++) {
      const int idx = tid + it * TAIL_THREADS;
      if (idx < nz) b.cz[idx] = vz[it];
      if (idx < nx) b.cx[idx] = vx[it];
    }
  }
  {
    const TailBufs<T> b = tail_bufs(base, a, 0);
#pragma unroll
    for (int it = 0; it < 2; it++) {
      const int idx = tid + it * TAIL_THREADS;
      if (idx < L0.N * L0.K) {
        b.u[idx] = T(0);
        b.f[idx] = fin[it] ? fv[it] : T(0);
      }
    }
  }
  __syncthreads();
  for (int l = 0; l < a.nlev; l++) {
    const TailLevel &L = a.lv[l];
    const T DH2 = (T)L.DH2;
    const TailBufs<T> Bl = tail_bufs(base, a, l);
    if (l == a.nlev - 1) {
      if (a.minv) {
        // e = B^{-1} (DH^2 f), rows summed in column order; Gs was filled by
        // the restriction onto this level unless it is the entry level
        const float *minv = reinterpret_cast<const float *>(a.minv);
        const int W = L.K - 2;
        if (l == 0) {
          for (int i = tid; i < a.cn; i += blockDim.x) {
            const int j = i / W, k = i - j * W + 1;
            Gs[i] = DH2 * Bl.f[j * L.K + k];
          }
          __syncthreads();
        }
        coarse_matvec<T>(minv, Gs, Ps, a.cn, Bl.u, W, L.K, (size_t)L.K);
      } else {
        const bool one_wave = L.N * L.K <= 512;
        tail_smooth<T, EXACT>(Bl, L.N, L.K, DH2, relax, a.nu_coarse, one_wave);
      }
      __syncthreads();
      break;
    }
    tail_smooth<T, EXACT>(Bl, L.N, L.K, DH2, relax, a.nu_pre, false);
    for (int idx = tid; idx < L.N * L.K; idx += blockDim.x) {
      const int j = idx / L.K, k = idx - j * L.K;
      T r = T(0);
      if (k >= 1 && k <= L.K - 2) {
        const int jS = (j == 0) ? 1 : j - 1;
        const int jN = (j == L.N - 1) ? L.N - 2 : j + 1;
        const T m1 = Bl.cx[idx], m2 = Bl.cx[idx + 1], m3 = Bl.cz[idx],
                m4 = Bl.cz[idx + L.K];
        const T *u = Bl.u;
        r = Bl.f[idx] - (m1 * u[idx - 1] + m2 * u[idx + 1] + m3 * u[jS * L.K + k] +
                           m4 * u[jN * L.K + k] - (m1 + m2 + m3 + m4) * u[idx]) /
                              DH2;
      }
      Rs[idx] = r;
    }
    __syncthreads();
    const TailLevel &C = a.lv[l + 1];
    const TailBufs<T> Bc = tail_bufs(base, a, l + 1);
    // onto the coarsest level of a direct solve: also its right-hand side
    const bool direct = a.minv && l + 1 == a.nlev - 1;
    const T DH2c = (T)C.DH2;
    for (int idx = tid; idx < C.N * C.K; idx += blockDim.x) {
      const int j = idx / C.K, k = idx - j * C.K;
      T v = T(0);
      if (k >= 1 && k <= C.K - 2) {
        const int jS = (j == 0) ? 1 : 2 * j - 1;
        const int jN = (j == C.N - 1) ? 2 * j - 1 : 2 * j + 1;
        const T *rc = Rs + (2 * j) * L.K;
        const T *rs = Rs + jS * L.K;
        const T *rn = Rs + jN * L.K;
        const int kf = 2 * k;
        v = T(0.0625) *
            (T(4.0) * rc[kf] + T(2.0) * rc[kf - 1] + T(2.0) * rc[kf + 1] +
             T(2.0) * rs[kf] + T(2.0) * rn[kf] + rs[kf - 1] + rs[kf + 1] +
             rn[kf - 1] + rn[kf + 1]);
        if (direct) Gs[j * (C.K - 2) + (k - 1)] = DH2c * v;
      }
      Bc.f[idx] = v;
      Bc.u[idx] = T(0);
    }
    __syncthreads();
  }
  for (int l = a.nlev - 2; l >= 0; l--) {
    const TailLevel &L = a.lv[l];
    const TailLevel &C = a.lv[l + 1];
    const TailBufs<T> Bl = tail_bufs(base, a, l), Bc = tail_bufs(base, a, l + 1);
    const int jmax = FULL_TOP ? C.N : C.N - 1;
    for (int idx = tid; idx < jmax * (C.K - 1); idx += blockDim.x) {
      const int j = idx / (C.K - 1), k = idx - j * (C.K - 1);
      const T *ec = Bc.u + j * C.K + k;
      T *uf = Bl.u + (2 * j) * L.K + 2 * k;
      const T e00 = ec[0], e01 = ec[1];
      uf[0] += e00;
      uf[1] += T(0.5) * (e00 + e01);
      if (!FULL_TOP || j < C.N - 1) {
        const T e10 = ec[C.K], e11 = ec[C.K + 1];
        T *uf1 = uf + L.K;
        uf1[0] += T(0.5) * (e00 + e10);
        uf1[1] += T(0.25) * (e00 + e01 + e10 + e11);
      }
    }
    __syncthreads();
    tail_smooth<T, EXACT>(Bl, L.N, L.K, (T)L.DH2, relax, a.nu_post, false);
  }
  {
    const TailLevel &L0 = a.lv[0];
    T *eo = reinterpret_cast<T *>(a.e_out);
    const T *u0 = tail_bufs(base, a, 0).u;
    for (int idx = tid; idx < L0.N * L0.K; idx += blockDim.x) {
      const int j = idx / L0.K, k = idx - j * L0.K;
      eo[(size_t)j * L0.P + k] = u0[idx];
    }
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

static size_t tile_lds_bytes(int TX, int TY, int H, bool has_f, size_t esz,
                             bool down, int threads) {
  const size_t R = TY + 2 * H, C = TX + 2 * H;
  size_t words = 3 * R * C + C;            // su, scx, scz (+1 row)
  if (has_f || (down && !res_in_regs(TX, TY, threads)))
    words += R * C;                        // sf, or residual scratch
  return words * esz;
}

// First level handled by the tail (or L if the coarsest level is too big).
int tail_entry_level(const qdeq_ctx *c) {
  for (int l = 0; l < c->L; l++)
    if (c->lv[l].N <= TAIL_MAX_N && c->lv[l].K <= TAIL_MAX_N && l > 0) {
      if (c->L - l > 6) return c->L;
      return l;
    }
  return c->L;
}

template <typename T>
static size_t tail_lds_bytes(const qdeq_ctx *c, int lt) {
  size_t n = 0;
  for (int l = lt; l < c->L; l++) n += tail_level_words(c->lv[l].N, c->lv[l].K);
  const size_t nr = (size_t)c->lv[lt].N * c->lv[lt].K;
  n += nr > (size_t)c->coarse_n ? nr : (size_t)c->coarse_n;
  n += c->coarse_n;  // right-hand side of the direct coarsest solve
  n += 1024;         // its partial sums
  return n * sizeof(T);
}

bool tail_fits(const qdeq_ctx *c, bool fp64) {
  const int lt = tail_entry_level(c);
  if (lt >= c->L) return false;
  const size_t b = fp64 ? tail_lds_bytes<double>(c, lt) : tail_lds_bytes<float>(c, lt);
  return b + 4096 <= c->lds_max;
}

// Shape of the down (smoothing + restriction) or up (prolongation +
// smoothing) kernel of a level.
static TileShape tile_shape(const qdeq_ctx *c, int level, bool down) {
  if (c->tuned) return down ? c->tdown[level] : c->tup[level];
  return TileShape{c->tile_tx, c->tile_ty, TY_THREADS};
}

// Whether a shape fits the kernels: even sizes, tile rows (with halo) of at
// most 320 points, and the LDS with 8 KiB to spare.
static bool shape_fits(const qdeq_ctx *c, const TileShape &s, bool down,
                       bool has_f, size_t esz) {
  const int nu = down ? c->cfg.n_pre : c->cfg.n_post;
  const int H = down ? 2 * nu + 2 : 2 * nu + 1;
  if (s.tx < 2 || s.ty < 2 || (s.tx & 1) || (s.ty & 1)) return false;
  if (s.tyt < 1 || s.tyt > TY_THREADS || (s.tyt & (s.tyt - 1))) return false;
  if (s.tx + 2 * H > 5 * TX_THREADS) return false;
  return tile_lds_bytes(s.tx, s.ty, H, has_f, esz, down, TX_THREADS * s.tyt) + 8192 <=
         c->lds_max;
}

// Largest dynamic LDS the tile kernels of this context would request.
bool tiles_fit(const qdeq_ctx *c, int tx, int ty) {
  const bool fp64 = c->cfg.precision != QDEQ_PREC_FP32_IR;
  const size_t esz = fp64 ? 8 : 4;
  const size_t down = tile_lds_bytes(tx, ty, 2 * c->cfg.n_pre + 2, true, esz, true,
                                     TILE_THREADS);
  const size_t up = tile_lds_bytes(tx, ty, 2 * c->cfg.n_post + 1, true, esz, false,
                                   TILE_THREADS);
  // tile rows (with halo) of at most 320 points
  if (tx + 2 * (2 * c->cfg.n_pre + 2) > 5 * TX_THREADS) return false;
  if (tx + 2 * (2 * c->cfg.n_post + 1) > 5 * TX_THREADS) return false;
  return down + 8192 <= c->lds_max && up + 8192 <= c->lds_max;
}

template <typename T, bool EXACT>
static void launch_down(qdeq_ctx *c, LevelBufs<T> *B, int l, bool has_f,
                        bool zero_u, int smode = DOWN_PLAIN) {
  const LevelDims &d = c->lv[l];
  const LevelDims &dc = c->lv[l + 1];
  const TileShape sh = tile_shape(c, l, true);
  const int TX = sh.tx, TY = sh.ty;
  const int nu = c->cfg.n_pre;
  const int H = 2 * nu + 2;
  const dim3 grd((d.K + TX - 1) / TX, (d.N + TY - 1) / TY);
  const dim3 blk(TX_THREADS, sh.tyt);
  const size_t lds = tile_lds_bytes(TX, TY, H, has_f, sizeof(T), true,
                                    TX_THREADS * sh.tyt);
  const T DH2 = (T)(d.DH * d.DH), relax = (T)c->cfg.relax;
  StartArgs sa{};
  if (smode != DOWN_PLAIN) {
    const TileShape s0 = tile_shape(c, 0, true);
    const LevelDims &d0 = c->lv[0];
    sa.part = c->partial + kStartOff;
    sa.nb = ((d0.K + s0.tx - 1) / s0.tx) * ((d0.N + s0.ty - 1) / s0.ty);
    sa.scal = c->scal;
    sa.mb = mb_dev(c);
    sa.gate = c->gate;
    sa.spec = c->spec_launch ? c->spec_go : 0;
    sa.rel = c->cfg.tol_mode == QDEQ_TOL_REL ? 1 : 0;
    sa.cap = c->cfg.max_vcycles;
    sa.tol = c->cfg.tol;
  }
  if (smode == DOWN_PLAIN && l == 0 && !has_f && c->fold_next && c->gate_ptr) {
    const TileShape u0 = tile_shape(c, 0, false);
    sa.fold = 1;
    sa.part = c->partial;
    sa.nb = ((d.K + u0.tx - 1) / u0.tx) * ((d.N + u0.ty - 1) / u0.ty);
    sa.scal = c->scal;
    sa.mb = mb_dev(c);
    sa.gate = c->gate;
    sa.rel = c->cfg.tol_mode == QDEQ_TOL_REL ? 1 : 0;
    sa.cap = c->cfg.max_vcycles;
    sa.tol = c->cfg.tol;
  }
#define DOWN_ARGS                                                                   \
  grd, blk, lds, c->stream, B[l].u, B[l].r, B[l].f, B[l].cx, B[l].cz, d.N, d.K, d.P, \
      DH2, relax, nu, zero_u ? 1 : 0, B[l + 1].f, dc.N, dc.K, dc.P, TX, TY,          \
      (const int *)c->gate_ptr, sa
  if (smode == START_PARTIALS)
    hipLaunchKernelGGL((k_down<T, false, EXACT, START_PARTIALS>), DOWN_ARGS);
  else if (smode == START_TEST)
    hipLaunchKernelGGL((k_down<T, true, EXACT, START_TEST>), DOWN_ARGS);
  else if (has_f)
    hipLaunchKernelGGL((k_down<T, true, EXACT, DOWN_PLAIN>), DOWN_ARGS);
  else
    hipLaunchKernelGGL((k_down<T, false, EXACT, DOWN_PLAIN>), DOWN_ARGS);
#undef DOWN_ARGS
}

// Whether a batch launched kernel by kernel may fold its stop tests into the
// next V-cycle's level-0 down kernel: the FP64 tiled cycle, at most 256
// level-0 up tiles (k_norm_gate's threads each sum one partial) and a
// level-0 down workgroup of four wavefronts or more.
bool fold_fits(const qdeq_ctx *c) {
  if (c->fold_off || c->cfg.precision != QDEQ_PREC_FP64 || !use_tiled(c) ||
      !launch_direct(c) || tail_entry_level(c) < 1)
    return false;
  const LevelDims &d = c->lv[0];
  const TileShape dn = tile_shape(c, 0, true), up = tile_shape(c, 0, false);
  const long nb = (long)((d.K + up.tx - 1) / up.tx) * ((d.N + up.ty - 1) / up.ty);
  return nb <= 256 && TX_THREADS * dn.tyt >= 256;
}

// Whether the start of the next FP64 solve may be folded into its first
// V-cycle's down kernels: two tile levels at least, and the start
// partials of every level-0 tile and the up kernel's norm partials fit
// their regions of c->partial.
bool fused_start_fits(const qdeq_ctx *c) {
  if (c->cfg.precision != QDEQ_PREC_FP64 || !use_tiled(c) || tail_entry_level(c) < 2)
    return false;
  const LevelDims &d = c->lv[0];
  const TileShape dn = tile_shape(c, 0, true), up = tile_shape(c, 0, false);
  const long nd = (long)((d.K + dn.tx - 1) / dn.tx) * ((d.N + dn.ty - 1) / dn.ty);
  const long nu = (long)((d.K + up.tx - 1) / up.tx) * ((d.N + up.ty - 1) / up.ty);
  return nd <= kStartCap && nu <= kStartOff;
}

template <typename T, bool EXACT, bool HAS_F, int MODE>
static void launch_up_t(qdeq_ctx *c, LevelBufs<T> *B, int l, double *partial,
                        double *u64, int *nblocks) {
  const LevelDims &d = c->lv[l];
  const LevelDims &dc = c->lv[l + 1];
  const TileShape sh = tile_shape(c, l, false);
  const int TX = sh.tx, TY = sh.ty;
  const dim3 grd((d.K + TX - 1) / TX, (d.N + TY - 1) / TY);
  if (nblocks) *nblocks = (int)(grd.x * grd.y);
  const int nu = c->cfg.n_post;
  const int H = 2 * nu + (MODE == UP_NORM ? 1 : 0);
  const size_t lds = tile_lds_bytes(TX, TY, H, HAS_F, sizeof(T), false,
                                    TX_THREADS * sh.tyt);
  const T DH2 = (T)(d.DH * d.DH), relax = (T)c->cfg.relax;
  hipLaunchKernelGGL((k_up<T, HAS_F, EXACT, true, MODE>), grd, dim3(TX_THREADS, sh.tyt),
                     lds, c->stream, B[l].r, B[l].u, B[l].f, B[l].cx, B[l].cz,
                     d.N, d.K, d.P, DH2, relax, nu, B[l + 1].u, dc.N, dc.K, dc.P,
                     TX, TY, partial, u64, (const int *)c->gate_ptr);
}

template <typename T, bool EXACT>
static void launch_tail(qdeq_ctx *c, LevelBufs<T> *B, int lt) {
  TailArgs a{};
  a.nlev = c->L - lt;
  for (int i = 0; i < a.nlev; i++) {
    const LevelDims &d = c->lv[lt + i];
    a.lv[i] = TailLevel{B[lt + i].cx, B[lt + i].cz, d.N, d.K, d.P, d.DH * d.DH};
  }
  a.f_in = B[lt].f;
  a.e_out = B[lt].u;
  a.nu_pre = c->cfg.n_pre;
  a.nu_post = c->cfg.n_post;
  a.nu_coarse = c->cfg.n_coarse;
  a.relax = c->cfg.relax;
  a.cn = c->coarse_n;
  a.minv = c->coarse_n > 0 ? (const void *)c->cinv32 : nullptr;
  a.gate = c->gate_ptr;
  hipLaunchKernelGGL((k_tail<T, EXACT, true>), dim3(1), dim3(TAIL_THREADS),
                     tail_lds_bytes<T>(c, lt), c->stream, a);
}

// Fast V-cycle for the FP64 solve (level 0 = the displacement, f = 0) or
// the FP32 correction of iterative refinement (level 0 = correction e,
// f = r32; the last up kernel adds e into the FP64 displacement).  The
// level-0 norm partials land in `partial` (FP64) and the count of partials
// in *nparts.
template <typename T, bool EXACT>
static void vcycle_tiled(qdeq_ctx *c, LevelBufs<T> *B, bool level0_has_f,
                         double *partial, double *u64, int *nparts) {
  const int lt = tail_entry_level(c);
  // the start of the solve folded into this cycle (FP64 only)
  const bool fuse = c->fuse_start_next && !level0_has_f && lt >= 2;
  c->fuse_start_next = false;
  for (int l = 0; l < lt; l++) {
    if (l == 1) timing_mark(c, QDEQ_TIME_DOWN1, 0);
    const int smode = !fuse ? DOWN_PLAIN : (l == 0 ? START_PARTIALS
                                                    : (l == 1 ? START_TEST : DOWN_PLAIN));
    launch_down<T, EXACT>(c, B, l, l > 0 || level0_has_f, l > 0 || level0_has_f, smode);
    if (l == 1) timing_mark(c, QDEQ_TIME_DOWN1, 1);
  }
  timing_mark(c, QDEQ_TIME_TAIL, 0);
  launch_tail<T, EXACT>(c, B, lt);
  timing_mark(c, QDEQ_TIME_TAIL, 1);
  for (int l = lt - 1; l >= 1; l--)
    launch_up_t<T, EXACT, true, UP_STORE>(c, B, l, nullptr, nullptr, nullptr);
  timing_mark(c, QDEQ_TIME_UP0, 0);
  if (u64)
    launch_up_t<T, EXACT, true, UP_ADD64>(c, B, 0, nullptr, u64, nullptr);
  else if (level0_has_f)
    launch_up_t<T, EXACT, true, UP_STORE>(c, B, 0, nullptr, nullptr, nullptr);
  else
    launch_up_t<T, EXACT, false, UP_NORM>(c, B, 0, partial, nullptr, nparts);
  timing_mark(c, QDEQ_TIME_UP0, 1);
}

void vcycle_tiled_fp64(qdeq_ctx *c, double *partial, int *nparts) {
  vcycle_tiled<double, true>(c, c->d, false, partial, nullptr, nparts);
}

void vcycle_tiled_fp32_ir(qdeq_ctx *c) {
  vcycle_tiled<float, false>(c, c->s, true, nullptr, c->d[0].u, nullptr);
}

// ---- tile-shape autotuner ----

// Launches one tile kernel of level l (down or up, in the solve's mode).
template <typename T, bool EXACT>
static void launch_one(qdeq_ctx *c, LevelBufs<T> *B, int l, bool down,
                       bool level0_has_f, double *u64) {
  const bool hf = l > 0 || level0_has_f;
  if (down) {
    launch_down<T, EXACT>(c, B, l, hf, hf);
  } else if (l > 0) {
    launch_up_t<T, EXACT, true, UP_STORE>(c, B, l, nullptr, nullptr, nullptr);
  } else if (u64) {
    launch_up_t<T, EXACT, true, UP_ADD64>(c, B, 0, nullptr, u64, nullptr);
  } else {
    int nb = 0;
    launch_up_t<T, EXACT, false, UP_NORM>(c, B, 0, c->partial, nullptr, &nb);
  }
}

static const int kTuneTX[] = {32, 64, 96, 128, 192, 256};
static const int kTuneTY[] = {2, 4, 8, 16, 32, 64};
static const int kTuneTYT[] = {4, 8, 16};
// Tile sizes that split a level into a whole number of tiles: the grids are
// 2^k + 1 points wide, so the power-of-two sizes leave a last tile of one
// or two points.  With one tile per CU per round, a grid of 16 x 16 tiles
// of 66 points at 1025^2 runs in one round where 96 x 32 needs two.
static const int kFitTX[] = {4, 8, 12, 16};
static const int kFitTY[] = {8, 12, 16, 24, 32};

static void add_fitted(std::vector<int> &v, int n, const int *parts, int np) {
  for (int i = 0; i < np; i++) {
    int t = (n + parts[i] - 1) / parts[i];
    t += t & 1;
    if (t >= 2 && std::find(v.begin(), v.end(), t) == v.end()) v.push_back(t);
  }
}

template <typename T, bool EXACT>
static int autotune_t(qdeq_ctx *c, LevelBufs<T> *B, bool level0_has_f,
                      double *u64, int reps) {
  const int lt = tail_entry_level(c);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  QDEQ_HIP(c, hipEventCreate(&e0));
  QDEQ_HIP(c, hipEventCreate(&e1));
  int st = QDEQ_OK;
  for (int l = 0; l < lt && st == QDEQ_OK; l++) {
    for (int dir = 0; dir < 2 && st == QDEQ_OK; dir++) {
      const bool down = dir == 0;
      const bool hf = l > 0 || level0_has_f;
      // every candidate's time, in a fixed order; the first one within 3%
      // of the fastest wins (stable against timing noise between runs)
      std::vector<std::pair<TileShape, float>> times;
      std::vector<int> txs(std::begin(kTuneTX), std::end(kTuneTX));
      std::vector<int> tys(std::begin(kTuneTY), std::end(kTuneTY));
      add_fitted(txs, c->lv[l].K, kFitTX, (int)(sizeof(kFitTX) / sizeof(int)));
      add_fitted(tys, c->lv[l].N, kFitTY, (int)(sizeof(kFitTY) / sizeof(int)));
      for (int tx : txs)
        for (int ty : tys)
          for (int tyt : kTuneTYT) {
            const TileShape cand{tx, ty, tyt};
            if (!shape_fits(c, cand, down, hf, sizeof(T))) continue;
            if (tyt > ty + 8) continue;  // more thread rows than tile rows
            const LevelDims &d = c->lv[l];
            const long blocks = (long)((d.K + tx - 1) / tx) * ((d.N + ty - 1) / ty);
            if (!down && l == 0 && !u64 && blocks > kPartialCap) continue;
            (down ? c->tdown[l] : c->tup[l]) = cand;
            launch_one<T, EXACT>(c, B, l, down, level0_has_f, u64);  // warm-up
            if (hipEventRecord(e0, c->stream) != hipSuccess) { st = QDEQ_ERR_HIP; break; }
            for (int r = 0; r < reps; r++)
              launch_one<T, EXACT>(c, B, l, down, level0_has_f, u64);
            if (hipEventRecord(e1, c->stream) != hipSuccess ||
                hipEventSynchronize(e1) != hipSuccess) { st = QDEQ_ERR_HIP; break; }
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { st = QDEQ_ERR_HIP; break; }
            times.emplace_back(cand, ms);
          }
      if (times.empty()) continue;
      float best_ms = times[0].second;
      for (const auto &t : times) best_ms = std::min(best_ms, t.second);
      for (const auto &t : times)
        if (t.second <= 1.03f * best_ms) {
          (down ? c->tdown[l] : c->tup[l]) = t.first;
          break;
        }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st != QDEQ_OK)
    return set_error(c, st, std::string("autotune: ") + hipGetErrorString(hipGetLastError()));
  QDEQ_HIP(c, hipGetLastError());
  return QDEQ_OK;
}

// Sets the per-level shapes explicitly (validated like tuned ones).
int set_tile_shapes(qdeq_ctx *c, int nlevels, const int *shapes) {
  if (!use_tiled(c))
    return set_error(c, QDEQ_ERR_STATE, "tile shapes: the tiled fast cycle is not in use");
  const int lt = tail_entry_level(c);
  if (nlevels != lt)
    return set_error(c, QDEQ_ERR_ARG, "tile shapes: need one entry per tile level (" +
                                          std::to_string(lt) + ")");
  const bool ir = c->cfg.precision == QDEQ_PREC_FP32_IR;
  const size_t esz = ir ? 4 : 8;
  TileShape dn[QDEQ_MAX_LEVELS], up[QDEQ_MAX_LEVELS];
  for (int l = 0; l < lt; l++) {
    dn[l] = TileShape{shapes[6 * l], shapes[6 * l + 1], shapes[6 * l + 2]};
    up[l] = TileShape{shapes[6 * l + 3], shapes[6 * l + 4], shapes[6 * l + 5]};
    const bool hf = l > 0 || ir;
    if (!shape_fits(c, dn[l], true, hf, esz) || !shape_fits(c, up[l], false, hf, esz))
      return set_error(c, QDEQ_ERR_ARG, "tile shapes: level " + std::to_string(l) +
                                            " does not fit the kernels");
    const LevelDims &d = c->lv[l];
    const long blocks = (long)((d.K + up[l].tx - 1) / up[l].tx) *
                        ((d.N + up[l].ty - 1) / up[l].ty);
    if (l == 0 && !ir && blocks > kPartialCap)
      return set_error(c, QDEQ_ERR_ARG, "tile shapes: too many level-0 blocks");
  }
  for (int l = 0; l < lt; l++) {
    c->tdown[l] = dn[l];
    c->tup[l] = up[l];
  }
  c->tuned = true;
  mg_destroy_graphs(c);
  return QDEQ_OK;
}

// Times every candidate shape of each tile level's down and up kernel on the
// context's own buffers and keeps the fastest.  Any shape gives the same
// results; the level-0 displacement is saved and restored around the runs.
int autotune_tiles(qdeq_ctx *c, int reps) {
  if (!use_tiled(c))
    return set_error(c, QDEQ_ERR_STATE, "autotune: the tiled fast cycle is not in use");
  const LevelDims &d0 = c->lv[0];
  const size_t n0 = (size_t)d0.N * d0.P;
  if ((size_t)2 * c->nx * c->nz < n0)
    return set_error(c, QDEQ_ERR_STATE, "autotune: staging buffer too small");
  QDEQ_HIP(c, hipMemcpyAsync(c->stage_U, c->d[0].u, n0 * sizeof(double),
                             hipMemcpyDeviceToDevice, c->stream));
  // start from the current default shapes
  for (int l = 0; l < QDEQ_MAX_LEVELS; l++) {
    c->tdown[l] = c->tup[l] = TileShape{c->tile_tx, c->tile_ty, TY_THREADS};
  }
  c->tuned = true;
  int st;
  if (c->cfg.precision == QDEQ_PREC_FP32_IR)
    st = autotune_t<float, false>(c, c->s, true, c->d[0].u, reps);
  else
    st = autotune_t<double, true>(c, c->d, false, nullptr, reps);
  QDEQ_HIP(c, hipMemcpyAsync(c->d[0].u, c->stage_U, n0 * sizeof(double),
                             hipMemcpyDeviceToDevice, c->stream));
  QDEQ_HIP(c, hipStreamSynchronize(c->stream));
  mg_destroy_graphs(c);
  if (st != QDEQ_OK) c->tuned = false;
  return st;
}

// Average duration of each kernel of the tiled cycle launched `reps` times
// back to back in its current shape: us[0 .. lt-1] = down kernels of levels
// 0 .. lt-1, us[lt] = tail, us[lt+1 ..] = up kernels of levels lt-1 .. 0.
template <typename T, bool EXACT>
static int breakdown_t(qdeq_ctx *c, LevelBufs<T> *B, bool level0_has_f,
                       double *u64, int reps, double *us) {
  const int lt = tail_entry_level(c);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  QDEQ_HIP(c, hipEventCreate(&e0));
  QDEQ_HIP(c, hipEventCreate(&e1));
  int st = QDEQ_OK;
  for (int i = 0; i < 2 * lt + 1 && st == QDEQ_OK; i++) {
    auto launch = [&]() {
      if (i < lt)
        launch_one<T, EXACT>(c, B, i, true, level0_has_f, u64);
      else if (i == lt)
        launch_tail<T, EXACT>(c, B, lt);
      else
        launch_one<T, EXACT>(c, B, 2 * lt - i, false, level0_has_f, u64);
    };
    launch();
    float ms = 0.f;
    if (hipEventRecord(e0, c->stream) != hipSuccess) st = QDEQ_ERR_HIP;
    for (int r = 0; r < reps && st == QDEQ_OK; r++) launch();
    if (st == QDEQ_OK &&
        (hipEventRecord(e1, c->stream) != hipSuccess ||
         hipEventSynchronize(e1) != hipSuccess ||
         hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
      st = QDEQ_ERR_HIP;
    us[i] = 1e3 * (double)ms / reps;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st != QDEQ_OK) return set_error(c, st, "kernel breakdown: HIP error");
  QDEQ_HIP(c, hipGetLastError());
  return QDEQ_OK;
}

int kernel_breakdown(qdeq_ctx *c, int reps, int *n, double *us) {
  if (!use_tiled(c))
    return set_error(c, QDEQ_ERR_STATE, "breakdown: the tiled fast cycle is not in use");
  const LevelDims &d0 = c->lv[0];
  const size_t n0 = (size_t)d0.N * d0.P;
  if ((size_t)2 * c->nx * c->nz < n0)
    return set_error(c, QDEQ_ERR_STATE, "breakdown: staging buffer too small");
  QDEQ_HIP(c, hipMemcpyAsync(c->stage_U, c->d[0].u, n0 * sizeof(double),
                             hipMemcpyDeviceToDevice, c->stream));
  *n = 2 * tail_entry_level(c) + 1;
  int st;
  if (c->cfg.precision == QDEQ_PREC_FP32_IR)
    st = breakdown_t<float, false>(c, c->s, true, c->d[0].u, reps, us);
  else
    st = breakdown_t<double, true>(c, c->d, false, nullptr, reps, us);
  QDEQ_HIP(c, hipMemcpyAsync(c->d[0].u, c->stage_U, n0 * sizeof(double),
                             hipMemcpyDeviceToDevice, c->stream));
  QDEQ_HIP(c, hipStreamSynchronize(c->stream));
  return st;
}

int tiled_setup(qdeq_ctx *c) {
  // Per-workgroup LDS available to dynamic allocations on this device; the
  // attribute raises the default cap where the runtime needs it.
  int dev = 0, optin = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipGetDeviceProperties(&prop, dev) == hipSuccess)
    c->lds_max = prop.sharedMemPerBlock;
  if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin,
                            dev) == hipSuccess &&
      (size_t)optin > c->lds_max)
    c->lds_max = (size_t)optin;
  const int big = (int)c->lds_max;
#define SETA(kern)                                                       \
  (void)hipFuncSetAttribute((const void *)(kern),                        \
                            hipFuncAttributeMaxDynamicSharedMemorySize, big)
  SETA((k_down<double, true, true, DOWN_PLAIN>));
  SETA((k_down<double, false, true, DOWN_PLAIN>));
  SETA((k_down<double, false, true, START_PARTIALS>));
  SETA((k_down<double, true, true, START_TEST>));
  SETA((k_down<float, true, false, DOWN_PLAIN>));
  SETA((k_down<float, false, false, DOWN_PLAIN>));
  SETA((k_up<double, true, true, true, UP_STORE>));
  SETA((k_up<double, false, true, true, UP_NORM>));
  SETA((k_up<float, true, false, true, UP_STORE>));
  SETA((k_up<float, true, false, true, UP_ADD64>));
  SETA((k_tail<double, true, true>));
  SETA((k_tail<float, false, true>));
#undef SETA
  (void)hipGetLastError();
  if (!tiles_fit(c, c->tile_tx, c->tile_ty)) c->tiled = false;
  return QDEQ_OK;
}

}  // namespace qdeq
