"""Synthetic QP batches for the parity tests and the benchmark (test data, not part of the library).

Every generator draws from ``numpy.random.default_rng(seed)`` in float64 and returns numpy
arrays; callers cast to float32 for single-precision runs.  Batched draws (one array per operand,
leading axis B) keep generation of large batches fast.

* :func:`random_feasible` -- dense QPs with a strictly feasible point:
  ``Q = M M^T + 1e-3 I``, ``q, A, G ~ N(0,1)``, ``x0 ~ N(0,1)``, ``b = A x0``,
  ``h = G x0 + |N(0,1)| + 0.1``.
* :func:`nnls` -- nonnegative least squares as a QP (the batch example of qpax,
  examples/batch_qp.py:14-33): ``Q = F^T F``, ``q = -F^T g``, ``G = -I``, ``h = 0``.
* :func:`planted_kkt` -- a QP built around a chosen KKT point, so its optimum is known in closed
  form (test/misc_test_utils.py:5-29).
* :func:`uneven` -- a mixed-size list of random feasible QPs.
"""

from __future__ import annotations

import numpy as np


def random_feasible(B, n, me, mi, seed=0):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((B, n, n))
    Q = M @ np.swapaxes(M, -1, -2) + 1e-3 * np.eye(n)
    Q = 0.5 * (Q + np.swapaxes(Q, -1, -2))
    q = rng.standard_normal((B, n))
    A = rng.standard_normal((B, me, n))
    G = rng.standard_normal((B, mi, n))
    x0 = rng.standard_normal((B, n))
    b = np.einsum("bij,bj->bi", A, x0)
    h = np.einsum("bij,bj->bi", G, x0) + np.abs(rng.standard_normal((B, mi))) + 0.1
    return Q, q, A, b, G, h


def nnls(B, n=5, m=10, seed=1):
    rng = np.random.default_rng(seed)
    F = rng.standard_normal((B, m, n))
    g = rng.standard_normal((B, m))
    Q = np.swapaxes(F, -1, -2) @ F
    q = -np.einsum("bji,bj->bi", F, g)
    G = np.broadcast_to(-np.eye(n), (B, n, n)).copy()
    h = np.zeros((B, n))
    A = np.zeros((B, 0, n))
    b = np.zeros((B, 0))
    return Q, q, A, b, G, h


def planted_kkt(B, nx, ns, ny, seed=0):
    """Returns (Q, q, A, b, G, h, x, s, z, y) with (x, s, z, y) an exact strictly complementary
    KKT point."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((B, nx))
    M = rng.standard_normal((B, nx, nx))
    Q = np.swapaxes(M, -1, -2) @ M
    G = rng.standard_normal((B, ns, nx))
    A = rng.standard_normal((B, ny, nx))
    s = np.abs(rng.standard_normal((B, ns)))
    z = np.abs(rng.standard_normal((B, ns)))
    coin = rng.random((B, ns)) < 0.5
    s = np.where(coin, 0.0, s)
    z = np.where(coin, z, 0.0)
    h = np.einsum("bij,bj->bi", G, x) + s
    b = np.einsum("bij,bj->bi", A, x)
    y = rng.standard_normal((B, ny))
    q = -np.einsum("bij,bj->bi", Q, x) - np.einsum("bji,bj->bi", G, z) - np.einsum("bji,bj->bi", A, y)
    return Q, q, A, b, G, h, x, s, z, y


def planted_kkt_legacy(nx, ns, ny, rs):
    """One problem drawn exactly as test/misc_test_utils.py:5-29 draws it from the legacy numpy
    generator ``rs`` (a numpy.random.RandomState), so a seeded sequence reproduces the
    reference's test problems.  Returns (Q, q, A, b, G, h, x, s, z, y)."""
    x = rs.randn(nx)
    M = rs.randn(nx, nx)
    Q = M.T @ M
    G = rs.randn(ns, nx)
    A = rs.randn(ny, nx)
    s = np.abs(rs.randn(ns))
    z = np.abs(rs.randn(ns))
    for i in range(ns):
        if rs.rand() < 0.5:
            s[i] = 0
        else:
            z[i] = 0
    h = G @ x + s
    b = A @ x
    y = rs.randn(ny)
    q = -Q @ x - G.T @ z - A.T @ y
    return Q, q, A, b, G, h, x, s, z, y


def uneven(B, seed=4, n_lo=8, n_hi=128):
    """B problems with n ~ U{n_lo..n_hi}, m_ineq ~ U{n..3n}, m_eq ~ U{0..n//4}.  Returns a list
    of (Q, q, A, b, G, h) tuples of unbatched arrays."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(B):
        n = int(rng.integers(n_lo, n_hi + 1))
        mi = int(rng.integers(n, 3 * n + 1))
        me = int(rng.integers(0, n // 4 + 1))
        sub = int(rng.integers(0, 2**31 - 1))
        Q, q, A, b, G, h = random_feasible(1, n, me, mi, seed=sub)
        out.append((Q[0], q[0], A[0], b[0], G[0], h[0]))
    return out


def elastic_infeasible(B, n, mi, frac=0.2, seed=5, me=0):
    """Random feasible QPs (``me`` equality rows) where a fraction ``frac`` of problems get
    ``h -= 3|N(0,1)|``, which typically makes them infeasible.  Returns (Q, q, G, h, shifted), or
    (Q, q, A, b, G, h, shifted) when ``me`` > 0."""
    Q, q, A, b, G, h = random_feasible(B, n, me, mi, seed=seed)
    rng = np.random.default_rng(seed + 1000)
    shifted = rng.random(B) < frac
    h = h - np.where(shifted[:, None], 3.0 * np.abs(rng.standard_normal((B, mi))), 0.0)
    if me:
        return Q, q, A, b, G, h, shifted
    return Q, q, G, h, shifted


def box_qp(B, n, me, dense, seed=0, inf_frac=0.0):
    """Random QPs with bounds lb <= x <= ub below `dense` general rows, laid out as the reference's
    test_QPTEST does (test/test_constraint_sizes.py:135-163): G = [G_dense; I; -I],
    h = [G_dense x0 + |N(0,1)| + 0.1; ub; -lb] with x0 ~ U(-0.5, 0.5) strictly inside,
    lb = -1 - U(0,1), ub = 1 + U(0,1), and a fraction inf_frac of the upper bounds set to +inf
    (rows the solver drops, pdip.py:166-178).  Q = M M^T + 1e-3 I, q, A ~ N(0,1), b = A x0."""
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((B, n, n))
    Q = M @ np.swapaxes(M, -1, -2) + 1e-3 * np.eye(n)
    Q = 0.5 * (Q + np.swapaxes(Q, -1, -2))
    q = rng.standard_normal((B, n))
    A = rng.standard_normal((B, me, n))
    x0 = rng.uniform(-0.5, 0.5, (B, n))
    b = np.einsum("bij,bj->bi", A, x0)
    Gd = rng.standard_normal((B, dense, n))
    hd = np.einsum("bij,bj->bi", Gd, x0) + np.abs(rng.standard_normal((B, dense))) + 0.1
    lb = -1.0 - rng.uniform(0.0, 1.0, (B, n))
    ub = 1.0 + rng.uniform(0.0, 1.0, (B, n))
    ub = np.where(rng.random((B, n)) < inf_frac, np.inf, ub)
    eye = np.broadcast_to(np.eye(n), (B, n, n))
    G = np.concatenate([Gd, eye, -eye], axis=1)
    h = np.concatenate([hd, ub, -lb], axis=1)
    return Q, q, A, b, G, h
