"""CPU restatement of the projection start's least-squares fit (test
infrastructure only).

No reference counterpart: the reference starts each elliptic solve from U as
the previous solve left it (odeSolver.cpp:652-720).  The engine can instead
start from a combination of the last few solutions
(qdeqcycle_fdm_amd/csrc/proj.hip): with d = g* - g_0 (new fault column minus
the newest stored one) and D_i = g_i - g_0 for the older stored columns, it
finds c minimising ||d - sum_i c_i D_i||_2 by CholeskyQR applied twice to
the augmented matrix A = [D_1 .. D_{m-1} | d] in one 256-thread workgroup:
G1 = A^T A, R1 = chol(G1), Q = A R1^{-1}; G2 = Q^T Q, R2 = chol(G2);
R = R2 R1; then R_DD c = R_Dd and ||d - D c|| = R_dd.  A D column whose
part independent of the columns before it has at most 1e-12 of its squared
norm in the first factorisation (a norm ratio of 1e-6) is left out.

This module repeats that computation in float64 numpy with the kernel's
summation order, so the coefficients can be compared bit for bit:

* a Gram entry: thread t (0..255) adds A_a[j] * A_b[j] for j = t, t+256, ...
  from 0.0; the 256 partial sums are added pairwise, entry t += entry t+s
  for s = 128, 64, ..., 1;
* Cholesky column by column, R_ki = (G_ki - sum_{l<k} R_lk R_li) * w_k and
  p_i = G_ii - sum_{l<i} R_li R_li, terms subtracted in increasing l, with
  R_ii = sqrt(p_i) and w_i = 1 / R_ii; a column left out has a zero column
  of R and w = 0 (its row of R is then zero, and every term involving it);
* Q_i = (A_i - sum_{k<i} R_ki Q_k) * w_i, terms subtracted in increasing k;
* R_ij = sum_{k=i..j} R2_ik R1_kj from 0.0 in increasing k;
* the back substitution from the last D column to the first, each row
  subtracting R_ik c_k in increasing k (c = 0 for a column left out).

``lstsq_reference`` solves the same problem with numpy's SVD-based
least squares, an independent method that pins the restatement.
"""
from __future__ import annotations

import numpy as np

THREADS = 256
DROP2 = 1e-12


def block_sum(prod: np.ndarray) -> float:
    """The kernel's sum of the N-vector prod (products already formed)."""
    n = prod.shape[0]
    acc = np.zeros(THREADS)
    for e in range(0, n, THREADS):
        chunk = prod[e:e + THREADS]
        acc[:chunk.shape[0]] = acc[:chunk.shape[0]] + chunk
    s = THREADS // 2
    while s >= 1:
        acc[:s] = acc[:s] + acc[s:2 * s]
        s //= 2
    return float(acc[0])


def _gram(A: list) -> dict:
    m = len(A)
    return {(a, b): block_sum(A[a] * A[b]) for a in range(m) for b in range(a, m)}


def _cholesky(G: dict, m: int, keep: list, drop: bool):
    """Returns (R, w) with w_i = 1 / R_ii (0 for a column left out or a zero
    pivot); keep[m-1] (d) is never set, d is always factored."""
    R = np.zeros((m, m))
    w = np.zeros(m)
    for i in range(m):
        for k in range(i):
            s = G[(k, i)]
            for l in range(k):
                s = s - R[l, k] * R[l, i]
            R[k, i] = s * w[k]
        p = G[(i, i)]
        for l in range(i):
            p = p - R[l, i] * R[l, i]
        used = i == m - 1 or keep[i]
        dropped = (not used) or (i < m - 1 and drop and
                                 not (G[(i, i)] > 0.0 and p > DROP2 * G[(i, i)]))
        if dropped:
            if i < m - 1:
                keep[i] = False
            R[:i + 1, i] = 0.0
        else:
            R[i, i] = np.sqrt(p) if p > 0.0 else 0.0
            w[i] = 1.0 / R[i, i] if p > 0.0 else 0.0
    return R, w


def _solve_columns(A: list, R: np.ndarray, w: np.ndarray) -> list:
    m = len(A)
    Q = [None] * m
    for i in range(m):
        q = A[i].copy()
        for k in range(i):
            q = q - R[k, i] * Q[k]
        Q[i] = q * w[i]
    return Q


def fit(g_new: np.ndarray, g_old: np.ndarray):
    """Coefficients of the fit: returns (c (cnt,), ||misfit|| / ||d||,
    columns kept), c[0] = 0 and c[i] the weight of (Z_i - Z_0)."""
    g_new = np.asarray(g_new, dtype=np.float64)
    g_old = np.asarray(g_old, dtype=np.float64)
    m = g_old.shape[0]
    g0 = g_old[0]
    A = [g_old[i + 1] - g0 for i in range(m - 1)] + [g_new - g0]
    keep = [True] * (m - 1) + [False]
    G1 = _gram(A)
    R1, w1 = _cholesky(G1, m, keep, True)
    Q = _solve_columns(A, R1, w1)
    G2 = _gram(Q)
    R2, _ = _cholesky(G2, m, keep, False)
    R = np.zeros((m, m))
    for i in range(m):
        for j in range(i, m):
            s = 0.0
            for k in range(i, j + 1):
                s = s + R2[i, k] * R1[k, j]
            R[i, j] = s
    c = [0.0] * (m - 1)
    for i in range(m - 2, -1, -1):
        acc = R[i, m - 1]
        for k in range(i + 1, m - 1):
            acc = acc - R[i, k] * c[k]
        c[i] = acc / R[i, i] if keep[i] else 0.0
    kept = sum(1 for k in keep[:m - 1] if k)
    out = np.concatenate([[0.0], c])
    dd = G1[(m - 1, m - 1)]
    ratio = R[m - 1, m - 1] / np.sqrt(dd) if dd > 0.0 else 0.0
    return out, float(ratio), kept


def lstsq_reference(g_new: np.ndarray, g_old: np.ndarray) -> np.ndarray:
    """The same least-squares problem by numpy's SVD solver (c[0] = 0)."""
    g0 = g_old[0]
    A = (g_old[1:] - g0).T
    sol = np.linalg.lstsq(A, g_new - g0, rcond=None)[0]
    return np.concatenate([[0.0], sol])


def apply(ring: np.ndarray, rr: np.ndarray, wr: np.ndarray, u: np.ndarray,
          c: np.ndarray) -> np.ndarray:
    """The start field: ring (cnt, N, P) the stored solutions (newest
    first), rr (cnt,) their remote values (row 0 of the remote column), wr
    (N, P) the response to a unit remote change, u (N, P) the field as the
    stage left it (boundary columns hold the new data; column K-1 = P-1
    here is the remote column), c from ``fit``.  Interior columns:
    U_0 + sum_i c_i (U_i - U_0) + w wr with w = r* - r_0 - sum_i c_i
    (r_i - r_0), each sum formed in that order; columns 0 and K-1 from u."""
    K = u.shape[1]
    r0 = rr[0]
    w = u[0, K - 1] - r0
    for i in range(1, ring.shape[0]):
        w = w - c[i] * (rr[i] - r0)
    acc = ring[0].copy()
    for i in range(1, ring.shape[0]):
        acc = acc + c[i] * (ring[i] - ring[0])
    out = acc + w * wr
    out[:, 0] = u[:, 0]
    out[:, K - 1] = u[:, K - 1]
    return out
