"""CPU restatement of the projection start's least-squares fit (test
infrastructure only).

No reference counterpart: the reference starts each elliptic solve from U as
the previous solve left it (odeSolver.cpp:652-720).  The engine can instead
start from a combination of the last few solutions
(qdeqcycle_fdm_amd/csrc/proj.hip): with d = g* - g_0 (new fault column minus
the newest stored one) and D_i = g_i - g_0 for the older stored columns, it
finds c minimising ||d - sum_i c_i D_i||_2 by classical Gram-Schmidt applied
twice (CGS2) in one 256-thread workgroup, and drops a column whose norm
after orthogonalisation is at most 1e-10 of its norm before.

This module repeats that computation in float64 numpy with the kernel's
summation order, so the coefficients can be compared bit for bit:

* a dot product over N entries: thread t (0..255) sums entries
  j = t, t+256, t+512, ... in that order from 0.0; within each 64-lane
  wavefront lane l adds the value of lane l^m for m = 32, 16, ..., 1
  (every lane ends with the same sum); the four wavefront sums are added as
  ((w0 + w1) + w2) + w3;
* an update x <- x - sum_q v_q Q_q subtracts the terms in q order;
* the back substitution runs from the last kept column to the first, each
  row subtracting R[a][b] c_b for b = a+1.. in order, then dividing.

``lstsq_reference`` solves the same problem with numpy's SVD-based
least squares, an independent method that pins the restatement.
"""
from __future__ import annotations

import numpy as np

THREADS = 256
DROP = 1e-10


def block_sum(prod: np.ndarray) -> float:
    """The kernel's sum of the N-vector prod (products already formed)."""
    n = prod.shape[0]
    acc = np.zeros(THREADS)
    for e in range(0, n, THREADS):
        chunk = prod[e:e + THREADS]
        acc[:chunk.shape[0]] = acc[:chunk.shape[0]] + chunk
    lanes = np.arange(64)
    waves = []
    for w in range(4):
        x = acc[64 * w:64 * (w + 1)].copy()
        m = 32
        while m >= 1:
            x = x + x[lanes ^ m]
            m >>= 1
        waves.append(x[0])
    return float(((waves[0] + waves[1]) + waves[2]) + waves[3])


def _cgs2(Q: list, x: np.ndarray, want_norm: bool):
    """Two passes of classical Gram-Schmidt of x against the columns in Q
    (in place on x); returns (coefficients summed over the passes, norm of
    x before the first pass or None)."""
    r = [0.0] * len(Q)
    nrm0 = None
    for p in range(2):
        v = [block_sum(q * x) for q in Q]
        if want_norm and p == 0:
            nrm0 = float(np.sqrt(block_sum(x * x)))
        for q_vec, vq in zip(Q, v):
            x -= vq * q_vec
        r = v if p == 0 else [a + b for a, b in zip(r, v)]
    return r, nrm0


def fit(g_new: np.ndarray, g_old: np.ndarray):
    """Coefficients of the fit: returns (c (cnt,), ||misfit|| / ||d||,
    columns kept), c[0] = 0 and c[i] the weight of (Z_i - Z_0)."""
    g_new = np.asarray(g_new, dtype=np.float64)
    g_old = np.asarray(g_old, dtype=np.float64)
    cnt = g_old.shape[0]
    g0 = g_old[0]
    d = g_new - g0
    D = [g_old[i] - g0 for i in range(1, cnt)]
    Q, kept = [], []
    R = {}
    for i, x in enumerate(D, start=1):
        x = x.copy()
        r, nrm0 = _cgs2(Q, x, True)
        nrm = float(np.sqrt(block_sum(x * x)))
        if nrm > 0.0 and nrm > DROP * nrm0:
            a = len(Q)
            for k, rk in enumerate(r):
                R[(k, a)] = rk
            R[(a, a)] = nrm
            Q.append(x / nrm)
            kept.append(i)
    d = d.copy()
    b, nd0 = _cgs2(Q, d, True)
    nd = float(np.sqrt(block_sum(d * d)))
    nq = len(Q)
    c = [0.0] * nq
    for a in range(nq - 1, -1, -1):
        acc = b[a]
        for bb in range(a + 1, nq):
            acc = acc - R[(a, bb)] * c[bb]
        c[a] = acc / R[(a, a)]
    out = np.zeros(cnt)
    for a, i in enumerate(kept):
        out[i] = c[a]
    ratio = nd / nd0 if nd0 > 0.0 else 0.0
    return out, ratio, nq


def lstsq_reference(g_new: np.ndarray, g_old: np.ndarray) -> np.ndarray:
    """The same least-squares problem by numpy's SVD solver (c[0] = 0)."""
    g0 = g_old[0]
    A = (g_old[1:] - g0).T
    sol = np.linalg.lstsq(A, g_new - g0, rcond=None)[0]
    return np.concatenate([[0.0], sol])


def apply(z: np.ndarray, wr: np.ndarray, u: np.ndarray, c: np.ndarray) -> np.ndarray:
    """The start field: z (cnt, N, P) stored solutions with the remote part
    removed (newest first), wr (N, P) the unit remote response, u (N, P) the
    field as the stage left it (boundary columns hold the new data, the
    remote value of row 0 at column K-1 = P-1 here), c from ``fit``.
    Interior columns: z0 + sum_i c_i (z_i - z0) + u[0, K-1] wr, added in that
    order; columns 0 and K-1 from u."""
    K = u.shape[1]
    acc = z[0].copy()
    for i in range(1, z.shape[0]):
        acc = acc + c[i] * (z[i] - z[0])
    out = acc + u[0, K - 1] * wr
    out[:, 0] = u[:, 0]
    out[:, K - 1] = u[:, K - 1]
    return out
