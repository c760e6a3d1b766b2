"""CPU oracle: a batched numpy restatement of qpax's PDIP QP solve, its kappa-relaxation and its
implicit-differentiation backward, plus the elastic (soft-constraint) variant.

TEST INFRASTRUCTURE ONLY.  This module is the checker the HIP library is compared against.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import it,
and only to check or to time a CPU baseline; the product path (``qpax_amd``) never calls it.

Every function is batched over a leading axis ``B`` (the reference is written per problem and
batched with ``jax.vmap``).  ``jax.lax.while_loop`` under ``vmap`` keeps iterating until every
problem is finished and select-masks the carry of finished problems, so each problem's result is
its unbatched result; :func:`_masked_loop` reproduces exactly that.  Arithmetic is carried out in
the dtype of the inputs (callers pass float64 for the oracle proper).  Linear solves follow the
reference: Householder QR (``numpy.linalg.qr``, LAPACK geqrf/orgqr, the routine XLA's CPU QR also
lowers to) followed by an upper-triangular back substitution, as in ``qpax/pdip.py:10-12``.

Pinning (what checks each function; see tests/test_oracle.py):
  * initialize, solve_qp (expanded elastic problem), initialize_elastic, solve_qp_elastic:
    the reference fixtures test/problem_data_pos.npz and test/initial_values_pos.npz at the
    reference's own 1e-6 (test/test_elastic_qp.py:64-107).
  * solve_qp special branches and inf rows: the known answers of test/test_constraint_sizes.py.
  * solve_qp on generate_random_qp problems: the closed-form optimum the generator plants
    (test/misc_test_utils.py:5-29) and the KKT conditions (misc_test_utils.py:32-43);
    NNLS problems (config-1 shape) against scipy.optimize.nnls.
  * relax_qp: the relaxed KKT system s*z = kappa holds at the result (it is the unique central-path
    point); a brute-force dense Newton solve of the relaxed KKT system gives the same point.
  * diff_qp / elastic backward: central finite differences of the tightly-converged relaxed
    solution map, and the elastic == expanded-plain gradient identity
    (test/test_elastic_qp.py:113-144).
"""

from __future__ import annotations

import numpy as np

MAX_ITERS = 30  # qpax/pdip.py:232, pdip_relaxed.py:76, elastic_qp.py:195,325


# ----------------------------------------------------------------------------------------------
# dense helpers (reference L0 boundary: jnp.linalg.qr + jax.scipy.linalg.solve_triangular)
# ----------------------------------------------------------------------------------------------
def _mv(M, v):
    """Batched M @ v for M (B,r,c), v (B,c)."""
    return np.matmul(M, v[..., None])[..., 0]


def _mtv(M, v):
    """Batched M^T @ v for M (B,r,c), v (B,r)."""
    return np.matmul(np.swapaxes(M, -1, -2), v[..., None])[..., 0]


def _back_substitute(R, y):
    """Solve R x = y, R upper triangular (B,k,k); y (B,k) or (B,k,c).  Row-by-row from the bottom,
    the order of LAPACK trsv/trsm for an upper, non-transposed, non-unit triangle."""
    vec = y.ndim == R.ndim - 1
    Y = y[..., None] if vec else y
    k = R.shape[-1]
    X = np.zeros_like(Y)
    for i in range(k - 1, -1, -1):
        acc = Y[:, i, :] - np.einsum("bj,bjc->bc", R[:, i, i + 1:], X[:, i + 1:, :])
        X[:, i, :] = acc / R[:, i, i][:, None]
    return X[..., 0] if vec else X


def qr(M):
    """Reduced QR of a batch of square matrices (jnp.linalg.qr default mode)."""
    if M.shape[-1] == 0:
        return (M.copy(), M.copy())
    Qm, R = np.linalg.qr(M)
    return (Qm, R)


def qr_solve(qrf, rhs):
    """qpax/pdip.py:10-12 -- R^{-1} (Q^T rhs)."""
    Qm, R = qrf
    if R.shape[-1] == 0:
        return rhs.copy()
    if rhs.ndim == Qm.ndim:  # matrix right-hand side
        return _back_substitute(R, np.matmul(np.swapaxes(Qm, -1, -2), rhs))
    return _back_substitute(R, _mtv(Qm, rhs))


# ----------------------------------------------------------------------------------------------
# standard-form PDIP (qpax/pdip.py)
# ----------------------------------------------------------------------------------------------
def initialize(Q, q, A, b, G, h):
    """qpax/pdip.py:15-35 -- CVXOPT/CVXGEN initialisation from one reduced KKT solve."""
    H = Q + np.matmul(np.swapaxes(G, -1, -2), G)
    L_H = qr(H)
    F = np.matmul(A, qr_solve(L_H, np.swapaxes(A, -1, -2)))
    L_F = qr(F)

    r1 = -q + _mtv(G, h)
    y = qr_solve(L_F, _mv(A, qr_solve(L_H, r1)) - b)
    x = qr_solve(L_H, r1 - _mtv(A, y))
    z = _mv(G, x) - h

    alpha_p = -np.min(-z, axis=-1, keepdims=True)
    s = np.where(alpha_p < 0, -z, -z + (1 + alpha_p))
    alpha_d = -np.min(z, axis=-1, keepdims=True)
    z = np.where(alpha_d >= 0, z + (1 + alpha_d), z)
    return x, s, z, y


def factorize_kkt(Q, G, A, s, z):
    """qpax/pdip.py:38-46 -- D = z/s, H = Q + G^T D G, F = A H^{-1} A^T, both QR-factored."""
    P_inv_vec = z / s
    H = Q + np.matmul(np.swapaxes(G, -1, -2), G * P_inv_vec[..., None])
    L_H = qr(H)
    F = np.matmul(A, qr_solve(L_H, np.swapaxes(A, -1, -2)))
    L_F = qr(F)
    return P_inv_vec, L_H, L_F


def solve_kkt_rhs(Q, G, A, s, z, P_inv_vec, L_H, L_F, v1, v2, v3, v4):
    """qpax/pdip.py:49-63 -- Algorithm 1, the cvxopt reduction of the KKT system."""
    r2 = v3 - v2 / z
    p1 = v1 + _mtv(G, P_inv_vec * r2)
    dy = qr_solve(L_F, _mv(A, qr_solve(L_H, p1)) - v4)
    dx = qr_solve(L_H, p1 - _mtv(A, dy))
    ds = v3 - _mv(G, dx)
    dz = (v2 - z * ds) / s
    return dx, ds, dz, dy


def ort_linesearch(x, dx):
    """qpax/pdip.py:66-73 -- max alpha <= 1 with x + alpha dx >= 0 (per problem)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = np.where(dx < 0, -x / dx, np.inf)
    if ratio.shape[-1] == 0:
        return np.ones(ratio.shape[:-1], dtype=x.dtype)
    return np.minimum(1.0, np.min(ratio, axis=-1)).astype(x.dtype)


def centering_params(s, z, ds_a, dz_a):
    """qpax/pdip.py:76-84 -- mu = s.z/m and Mehrotra's sigma = (gap_aff/gap)^3."""
    sz = np.sum(s * z, axis=-1)
    mu = sz / s.shape[-1]
    alpha = np.minimum(ort_linesearch(s, ds_a), ort_linesearch(z, dz_a))[..., None]
    sigma = (np.sum((s + alpha * ds_a) * (z + alpha * dz_a), axis=-1) / sz) ** 3
    return sigma, mu


def _inf_norm(*vs):
    parts = [np.abs(v) for v in vs if v.shape[-1] > 0]
    return np.max(np.concatenate(parts, axis=-1), axis=-1)


def pdip_pc_step(Q, q, A, b, G, h, x, s, z, y, solver_tol):
    """qpax/pdip.py:87-119 -- one Mehrotra predictor-corrector step.  Returns the stepped iterate
    and the convergence flag of the *entry* iterate (the step is taken regardless)."""
    r1 = _mv(Q, x) + q + _mtv(A, y) + _mtv(G, z)
    r2 = s * z
    r3 = _mv(G, x) + s - h
    r4 = _mv(A, x) - b
    converged = np.where(_inf_norm(r1, r2, r3, r4) < solver_tol, 1, 0).astype(np.int32)

    P_inv_vec, L_H, L_F = factorize_kkt(Q, G, A, s, z)
    _, ds_a, dz_a, _ = solve_kkt_rhs(Q, G, A, s, z, P_inv_vec, L_H, L_F, -r1, -r2, -r3, -r4)

    sigma, mu = centering_params(s, z, ds_a, dz_a)
    r2 = r2 - ((sigma * mu)[..., None] - (ds_a * dz_a))
    dx, ds, dz, dy = solve_kkt_rhs(Q, G, A, s, z, P_inv_vec, L_H, L_F, -r1, -r2, -r3, -r4)

    alpha = (0.99 * np.minimum(ort_linesearch(s, ds), ort_linesearch(z, dz)))[..., None]
    return x + alpha * dx, s + alpha * ds, z + alpha * dz, y + alpha * dy, converged


def _masked_loop(step, state, consts, max_iters=MAX_ITERS):
    """jax.lax.while_loop(iter < 30 and converged == 0, body) under jax.vmap: the body runs on the
    problems whose predicate holds, the others keep their carry (pdip.py:228-238)."""
    B = state[0].shape[0]
    conv = np.zeros(B, np.int32)
    iters = np.zeros(B, np.int32)
    state = [np.array(v, copy=True) for v in state]
    while True:
        act = np.nonzero((iters < max_iters) & (conv == 0))[0]
        if act.size == 0:
            break
        with np.errstate(all="ignore"):
            out = step(*[c[act] for c in consts], *[v[act] for v in state])
        for v, nv in zip(state, out[:-1]):
            v[act] = nv
        conv[act] = out[-1]
        iters[act] += 1
    return state, conv, iters


def solve_eq_only(Q, q, A, b):
    """qpax/pdip.py:137-150 -- equality-only QP via the Schur complement (Boyd p.559)."""
    Q_f = qr(Q)
    Qinv_At = qr_solve(Q_f, np.swapaxes(A, -1, -2))
    Qinv_g = qr_solve(Q_f, q)
    S = -np.matmul(A, Qinv_At)
    S_f = qr(S)
    y = qr_solve(S_f, _mv(A, Qinv_g) + b)
    x = qr_solve(Q_f, -_mtv(A, y) - q)
    return x, y


def remove_inf_constraints(G, h):
    """qpax/pdip.py:166-178 -- zero the rows of G (and entries of h) where h is +-inf."""
    mask = np.isinf(h)
    h2 = np.where(mask, 0, h)
    G2 = G * (~mask)[..., None].astype(G.dtype)
    return G2, h2, mask


def _as_batch(Q, q, A, b, G, h):
    Q = np.asarray(Q)
    q = np.asarray(q)
    B = q.shape[0]
    n = q.shape[-1]
    A = np.asarray(A).reshape(B, -1, n)
    b = np.asarray(b).reshape(B, -1)
    G = np.asarray(G).reshape(B, -1, n)
    h = np.asarray(h).reshape(B, -1)
    return Q.reshape(B, n, n), q, A, b, G, h


def solve_qp(Q, q, A, b, G, h, solver_tol=1e-3):
    """qpax/pdip.py:181-251 -- batched solve_qp; returns (x, s, z, y, converged, iters)."""
    Q, q, A, b, G, h = _as_batch(Q, q, A, b, G, h)
    B = q.shape[0]
    dt = q.dtype
    if b.shape[-1] == 0 and h.shape[-1] == 0:
        x = np.linalg.solve(Q, -q[..., None])[..., 0]
        e = np.zeros((B, 0), dt)
        return x, e, e, e, np.ones(B, np.int32), np.zeros(B, np.int32)
    if h.shape[-1] == 0:
        x, y = solve_eq_only(Q, q, A, b)
        e = np.zeros((B, 0), dt)
        return x, e, e, y, np.ones(B, np.int32), np.zeros(B, np.int32)

    G, h, mask = remove_inf_constraints(G, h)
    Q = 0.5 * (Q + np.swapaxes(Q, -1, -2))
    x, s, z, y = initialize(Q, q, A, b, G, h)

    def step(Q, q, A, b, G, h, x, s, z, y):
        return pdip_pc_step(Q, q, A, b, G, h, x, s, z, y, solver_tol)

    (x, s, z, y), conv, iters = _masked_loop(step, [x, s, z, y], [Q, q, A, b, G, h])
    z = np.where(mask, 0, z)
    s = np.where(mask, np.inf, s)
    return x, s, z, y, conv, iters


def solve_qp_trace(Q, q, A, b, G, h, solver_tol=1e-3, debug_step=False):
    """solve_qp with the per-iteration table of solve_qp_debug (qpax/pdip.py:254-360): returns
    ((x, s, z, y, converged, iters), trace) with trace[:, k] = (|r1|, |r2|, |r3|, |r4|, alpha) of
    iteration k (infinity norms at the iteration's entry, |r4| = 0 without equality rows as the
    debug print's zeros(1), pdip.py:285-286) and NaN rows after the last iteration.  debug_step
    takes the debug loop's step rule 0.99 min(1, 0.99 ort(s), 0.99 ort(z)) (pdip.py:281) instead
    of solve_qp's 0.99 min(ort(s), ort(z)) (pdip.py:112)."""
    Q, q, A, b, G, h = _as_batch(Q, q, A, b, G, h)
    B = q.shape[0]
    G, h, mask = remove_inf_constraints(G, h)
    Q = 0.5 * (Q + np.swapaxes(Q, -1, -2))
    x, s, z, y = initialize(Q, q, A, b, G, h)
    trace = np.full((B, MAX_ITERS, 5), np.nan)

    def norm(v):
        return np.max(np.abs(v), axis=-1) if v.shape[-1] else np.zeros(v.shape[0])

    def step(Q, q, A, b, G, h, x, s, z, y):
        r1 = _mv(Q, x) + q + _mtv(A, y) + _mtv(G, z)
        r2 = s * z
        r3 = _mv(G, x) + s - h
        r4 = _mv(A, x) - b
        rows = np.stack([norm(r1), norm(r2), norm(r3), norm(r4)], axis=-1)
        converged = np.where(_inf_norm(r1, r2, r3, r4) < solver_tol, 1, 0).astype(np.int32)
        P_inv_vec, L_H, L_F = factorize_kkt(Q, G, A, s, z)
        _, ds_a, dz_a, _ = solve_kkt_rhs(Q, G, A, s, z, P_inv_vec, L_H, L_F, -r1, -r2, -r3, -r4)
        sigma, mu = centering_params(s, z, ds_a, dz_a)
        r2 = r2 - ((sigma * mu)[..., None] - (ds_a * dz_a))
        dx, ds, dz, dy = solve_kkt_rhs(Q, G, A, s, z, P_inv_vec, L_H, L_F, -r1, -r2, -r3, -r4)
        a_s, a_z = ort_linesearch(s, ds), ort_linesearch(z, dz)
        if debug_step:
            alpha = 0.99 * np.minimum(np.minimum(1.0, 0.99 * a_s), 0.99 * a_z)
        else:
            alpha = 0.99 * np.minimum(a_s, a_z)
        rows = np.concatenate([rows, alpha[..., None]], axis=-1)
        a = alpha[..., None]
        return x + a * dx, s + a * ds, z + a * dz, y + a * dy, rows, converged

    conv = np.zeros(B, np.int32)
    iters = np.zeros(B, np.int32)
    state = [x, s, z, y]
    consts = [Q, q, A, b, G, h]
    while True:
        act = np.nonzero((iters < MAX_ITERS) & (conv == 0))[0]
        if act.size == 0:
            break
        with np.errstate(all="ignore"):
            out = step(*[c[act] for c in consts], *[v[act] for v in state])
        for v, nv in zip(state, out[:4]):
            v[act] = nv
        trace[act, iters[act]] = out[4]
        conv[act] = out[5]
        iters[act] += 1
    x, s, z, y = state
    z = np.where(mask, 0, z)
    s = np.where(mask, np.inf, s)
    return (x, s, z, y, conv, iters), trace


# ----------------------------------------------------------------------------------------------
# relaxation and implicit backward (qpax/pdip_relaxed.py, qpax/diff_qp.py)
# ----------------------------------------------------------------------------------------------
def pdip_newton_step(Q, q, A, b, G, h, x, s, z, y, solver_tol, target_kappa):
    """qpax/pdip_relaxed.py:7-52 -- damped Newton step on the kappa-relaxed KKT system."""
    r1 = _mv(Q, x) + q + _mtv(A, y) + _mtv(G, z)
    r2 = s * z - target_kappa
    r3 = _mv(G, x) + s - h
    r4 = _mv(A, x) - b
    converged = np.where(_inf_norm(r1, r2, r3, r4) < solver_tol, 1, 0).astype(np.int32)

    P_inv_vec, L_H, L_F = factorize_kkt(Q, G, A, s, z)
    dx, ds, dz, dy = solve_kkt_rhs(Q, G, A, s, z, P_inv_vec, L_H, L_F, -r1, -r2, -r3, -r4)

    alpha = 0.99 * np.minimum(np.minimum(1.0, 0.99 * ort_linesearch(s, ds)), 0.99 * ort_linesearch(z, dz))
    alpha = alpha[..., None]
    return x + alpha * dx, s + alpha * ds, z + alpha * dz, y + alpha * dy, converged


def relax_qp(Q, q, A, b, G, h, x, s, z, y, solver_tol=1e-5, target_kappa=1e-5):
    """qpax/pdip_relaxed.py:70-103 -- up to 30 relaxed Newton steps from a PDIP solution."""
    Q, q, A, b, G, h = _as_batch(Q, q, A, b, G, h)

    def step(Q, q, A, b, G, h, x, s, z, y):
        return pdip_newton_step(Q, q, A, b, G, h, x, s, z, y, solver_tol, target_kappa)

    (x, s, z, y), conv, iters = _masked_loop(step, [x, s, z, y], [Q, q, A, b, G, h])
    return x, s, z, y, conv, iters


def optnet_derivatives(dz, dlam, dnu, z, lam, nu, exact_dG=False):
    """qpax/diff_qp.py:8-17 -- parameter cotangents from the adjoint KKT solution.

    The reference scales both terms of dl_dG by diag(lam) (diff_qp.py:11), which makes its second
    term lam^2 dz^T; the exact vector-Jacobian product of the relaxed solution map is
    diag(lam) dlam z^T + lam dz^T.  exact_dG=True returns the latter (checked against central
    finite differences); the default restates the reference."""
    outer = lambda u, v: u[..., :, None] * v[..., None, :]  # noqa: E731
    dl_dQ = 0.5 * (outer(dz, z) + outer(z, dz))
    dl_dA = outer(dnu, z) + outer(nu, dz)
    if exact_dG:
        dl_dG = lam[..., :, None] * outer(dlam, z) + outer(lam, dz)
    else:
        dl_dG = lam[..., :, None] * (outer(dlam, z) + outer(lam, dz))
    return dl_dQ, dz, dl_dA, -dnu, dl_dG, -lam * dlam


def diff_qp(Q, q, A, b, G, h, z, s, lam, nu, dl_dz, exact_dG=False):
    """qpax/diff_qp.py:20-46 -- factor at the relaxed point, solve with rhs (-dl/dx, 0, 0, 0)."""
    Q, q, A, b, G, h = _as_batch(Q, q, A, b, G, h)
    P_inv_vec, L_H, L_F = factorize_kkt(Q, G, A, s, lam)
    zs = np.zeros_like(h)
    dz, ds, dlam_tilde, dnu = solve_kkt_rhs(Q, G, A, s, lam, P_inv_vec, L_H, L_F, -dl_dz, zs, zs, np.zeros_like(b))
    dlam = dlam_tilde / lam
    return optnet_derivatives(dz, dlam, dnu, z, lam, nu, exact_dG)


def solve_qp_primal_fwd(Q, q, A, b, G, h, solver_tol=1e-5, target_kappa=1e-3):
    """qpax/diff_qp.py:61-71 -- forward of the custom_vjp: solve, then relax from the solution.
    Returns x, the saved residuals (relaxed point) and both (converged, iters) pairs."""
    x, s, z, y, c1, i1 = solve_qp(Q, q, A, b, G, h, solver_tol=solver_tol)
    xr, sr, zr, yr, c2, i2 = relax_qp(Q, q, A, b, G, h, x, s, z, y, solver_tol=solver_tol,
                                      target_kappa=target_kappa)
    return x, (xr, sr, zr, yr), (c1, i1, c2, i2)


def solve_qp_primal_bwd(Q, q, A, b, G, h, relaxed, dl_dx, exact_dG=False):
    """qpax/diff_qp.py:74-79 -- the six cotangents (Q, q, A, b, G, h)."""
    xr, sr, zr, yr = relaxed
    return diff_qp(Q, q, A, b, G, h, xr, sr, zr, yr, dl_dx, exact_dG)


# ----------------------------------------------------------------------------------------------
# elastic QP (qpax/elastic_qp.py)
# ----------------------------------------------------------------------------------------------
def _aug_solve(H, A, r1, r4):
    """Solve [[H, A^T], [A, 0]] [x; y] = [r1; r4] the reference way: QR of H and of the Schur
    complement F = A H^{-1} A^T (qpax/pdip.py:17-24).  With no rows in A this is qr_solve(H, r1)."""
    L_H = qr(H)
    if A is None or A.shape[-2] == 0:
        return qr_solve(L_H, r1), None, L_H, None
    F = np.matmul(A, qr_solve(L_H, np.swapaxes(A, -1, -2)))
    L_F = qr(F)
    y = qr_solve(L_F, _mv(A, qr_solve(L_H, r1)) - r4)
    x = qr_solve(L_H, r1 - _mtv(A, y))
    return x, y, L_H, L_F


def _aug_solve_factored(L_H, L_F, A, r1, r4):
    if L_F is None:
        return qr_solve(L_H, r1), None
    y = qr_solve(L_F, _mv(A, qr_solve(L_H, r1)) - r4)
    return qr_solve(L_H, r1 - _mtv(A, y)), y


# Equality rows for the elastic QP are an extension: qpax's elastic functions have no A, b
# (qpax/elastic_qp.py:184,421).  With A = None every function below restates the reference
# exactly; with equality rows the one n x n solve of Algorithm 5 becomes the reduced KKT solve of
# [[H, A^T], [A, 0]] (as the standard form does), stationarity gains A^T y and the residuals gain
# A x - b.  The extension is checked against the expanded standard-form problem.
def solve_init_elastic_ls(Q, q, G, h, penalty, A=None, b=None):
    """qpax/elastic_qp.py:9-24."""
    ns = h.shape[-1]
    pen = penalty if np.ndim(penalty) == 0 else np.asarray(penalty).reshape(-1, 1)
    r1 = -q
    r2 = pen * np.ones((h.shape[0], ns), h.dtype)
    r4 = h
    H = Q + 0.5 * np.matmul(np.swapaxes(G, -1, -2), G)
    x, y, _, _ = _aug_solve(H, A, r1 - 0.5 * _mtv(G, r2 - r4), b)
    z2 = 0.5 * (_mv(G, x) + r2 - r4)
    z1 = r2 - z2
    t = -z1
    return np.concatenate((x, t), -1), np.concatenate((z1, z2), -1), y


def initialize_elastic(Q, q, G, h, penalty, A=None, b=None):
    """qpax/elastic_qp.py:27-47 -- least-squares init plus the CVXOPT shift on stacked (z1, z2).
    Returns (x, t, s1, s2, z1, z2) (and y when equality rows are given)."""
    x_big, z_big, y = solve_init_elastic_ls(Q, q, G, h, penalty, A, b)
    alpha_p = -np.min(-z_big, axis=-1, keepdims=True)
    s_big = np.where(alpha_p < 0, -z_big, -z_big + (1 + alpha_p))
    alpha_d = -np.min(z_big, axis=-1, keepdims=True)
    z_big = np.where(alpha_d >= 0, z_big + (1 + alpha_d), z_big)
    nx = q.shape[-1]
    ns = h.shape[-1]
    out = (x_big[:, :nx], x_big[:, nx:], s_big[:, :ns], s_big[:, ns:], z_big[:, :ns], z_big[:, ns:])
    return out if y is None else (*out, y)


def solve_elastic_kkt_affine(s1, z1, s2, z2, Q, G, r1, r2, r3, r4, r5, r6, L_H=None, A=None, r7=None, L_F=None):
    """qpax/elastic_qp.py:50-96 -- Algorithm 5 (one n x n factorisation); with L_H given this is
    solve_elastic_kkt_cc (the corrector reuses the factor).  Returns dx, dt, ds1, ds2, dz1, dz2,
    (L_H, L_F) and dy (None without equality rows)."""
    a1 = s1 / z1
    a2 = s2 / z2
    w1 = r3 / z1
    w2 = r4 / z2
    p1 = r5 - r6 + w2 - w1 - a1 * r2
    a3 = a1 + a2
    rhs = r1 - _mtv(G, p1 / a3)
    if L_H is None:
        H = Q + np.matmul(np.swapaxes(G, -1, -2), G * (1 / a3)[..., None])
        dx, dy, L_H, L_F = _aug_solve(H, A, rhs, r7)
    else:
        dx, dy = _aug_solve_factored(L_H, L_F, A, rhs, r7)
    dz2 = (p1 + _mv(G, dx)) / a3
    dz1 = -r2 - dz2
    ds1 = (r3 - s1 * dz1) / z1
    ds2 = (r4 - s2 * dz2) / z2
    dt = ds1 - r5
    return dx, dt, ds1, ds2, dz1, dz2, (L_H, L_F), dy


def _elastic_res(Q, q, G, h, pen, x, t, s1, s2, z1, z2, kappa, sign_t, sign_r2, A, b, y):
    r1 = _mv(Q, x) + q + _mtv(G, z2)
    if A is not None:
        r1 = r1 + _mtv(A, y)
    r2 = sign_r2 * (-z1 - z2 + pen)
    r3 = s1 * z1 - kappa
    r4 = s2 * z2 - kappa
    r5 = sign_t * t + s1
    r6 = _mv(G, x) + sign_t * t + s2 - h
    r7 = None if A is None else _mv(A, x) - b
    return r1, r2, r3, r4, r5, r6, r7


def pdip_pc_step_elastic(Q, q, G, h, penalty, x, t, s1, s2, z1, z2, solver_tol, A=None, b=None, y=None):
    """qpax/elastic_qp.py:99-174 -- Algorithm 4, one predictor-corrector step."""
    pen = penalty[..., None] if np.ndim(penalty) else penalty
    r1, r2, r3, r4, r5, r6, r7 = _elastic_res(Q, q, G, h, pen, x, t, s1, s2, z1, z2, 0.0, -1.0, 1.0, A, b, y)
    norms = (r1, r2, r3, r4, r5, r6) if r7 is None else (r1, r2, r3, r4, r5, r6, r7)
    converged = np.where(_inf_norm(*norms) < solver_tol, 1, 0).astype(np.int32)
    m7 = None if r7 is None else -r7
    _, _, ds1_a, ds2_a, dz1_a, dz2_a, (L_H, L_F), _ = solve_elastic_kkt_affine(
        s1, z1, s2, z2, Q, G, -r1, -r2, -r3, -r4, -r5, -r6, A=A, r7=m7)
    s = np.concatenate((s1, s2), -1)
    z = np.concatenate((z1, z2), -1)
    sigma, mu = centering_params(s, z, np.concatenate((ds1_a, ds2_a), -1), np.concatenate((dz1_a, dz2_a), -1))
    sm = (sigma * mu)[..., None]
    r3 = r3 - (sm - ds1_a * dz1_a)
    r4 = r4 - (sm - ds2_a * dz2_a)
    dx, dt, ds1, ds2, dz1, dz2, _, dy = solve_elastic_kkt_affine(
        s1, z1, s2, z2, Q, G, -r1, -r2, -r3, -r4, -r5, -r6, L_H=L_H, A=A, r7=m7, L_F=L_F)
    ds = np.concatenate((ds1, ds2), -1)
    dz = np.concatenate((dz1, dz2), -1)
    alpha = 0.99 * np.minimum(np.minimum(1.0, 0.99 * ort_linesearch(s, ds)), 0.99 * ort_linesearch(z, dz))
    a = alpha[..., None]
    out = [x + a * dx, t + a * dt, s1 + a * ds1, s2 + a * ds2, z1 + a * dz1, z2 + a * dz2]
    if A is not None:
        out.append(y + a * dy)
    return (*out, converged)


def _eq_batch(A, b, B, n):
    if A is None:
        return None, None
    A = np.asarray(A).reshape(B, -1, n)
    b = np.asarray(b).reshape(B, -1)
    if A.shape[1] == 0:
        return None, None
    return A, b


def solve_qp_elastic(Q, q, G, h, penalty, solver_tol=1e-3, A=None, b=None):
    """qpax/elastic_qp.py:184-228 -> (x, t, s1, s2, z1, z2, converged, iters) (y is appended
    after z2 when equality rows are given)."""
    Q = np.asarray(Q)
    q = np.asarray(q)
    B, n = q.shape
    G = np.asarray(G).reshape(B, -1, n)
    h = np.asarray(h).reshape(B, -1)
    A, b = _eq_batch(A, b, B, n)
    pen = np.broadcast_to(np.asarray(penalty, q.dtype), (B,)).copy()
    Q = 0.5 * (Q + np.swapaxes(Q, -1, -2))
    init = initialize_elastic(Q, q, G, h, pen[:, None], A, b)

    if A is None:
        def step(Q, q, G, h, pen, *st):
            return pdip_pc_step_elastic(Q, q, G, h, pen, *st, solver_tol)
        st, conv, iters = _masked_loop(step, list(init), [Q, q, G, h, pen])
    else:
        def step(Q, q, G, h, pen, A, b, *st):
            return pdip_pc_step_elastic(Q, q, G, h, pen, *st[:6], solver_tol, A=A, b=b, y=st[6])
        st, conv, iters = _masked_loop(step, list(init), [Q, q, G, h, pen, A, b])
    return (*st, conv, iters)


def pdip_newton_step_elastic(Q, q, G, h, penalty, x, t, s1, s2, z1, z2, solver_tol, target_kappa,
                             consistent=False, A=None, b=None, y=None):
    """qpax/elastic_qp.py:235-316.  The reference's relaxed residuals flip the sign of t and of r2
    relative to the forward residuals (elastic_qp.py:257,260-261) while reusing the forward linear
    solve; ``consistent=False`` restates that exactly.  ``consistent=True`` uses the forward signs
    (elastic_qp.py:105-110) with kappa subtracted from r3, r4."""
    pen = penalty[..., None] if np.ndim(penalty) else penalty
    sign = -1.0 if consistent else 1.0
    r1, r2, r3, r4, r5, r6, r7 = _elastic_res(Q, q, G, h, pen, x, t, s1, s2, z1, z2, target_kappa, sign,
                                              1.0 if consistent else -1.0, A, b, y)
    norms = (r1, r2, r3, r4, r5, r6) if r7 is None else (r1, r2, r3, r4, r5, r6, r7)
    converged = np.where(_inf_norm(*norms) < solver_tol, 1, 0).astype(np.int32)
    m7 = None if r7 is None else -r7
    dx, dt, ds1, ds2, dz1, dz2, _, dy = solve_elastic_kkt_affine(
        s1, z1, s2, z2, Q, G, -r1, -r2, -r3, -r4, -r5, -r6, A=A, r7=m7)
    s = np.concatenate((s1, s2), -1)
    z = np.concatenate((z1, z2), -1)
    ds = np.concatenate((ds1, ds2), -1)
    dz = np.concatenate((dz1, dz2), -1)
    alpha = 0.99 * np.minimum(np.minimum(1.0, 0.99 * ort_linesearch(s, ds)), 0.99 * ort_linesearch(z, dz))
    a = alpha[..., None]
    out = [x + a * dx, t + a * dt, s1 + a * ds1, s2 + a * ds2, z1 + a * dz1, z2 + a * dz2]
    if A is not None:
        out.append(y + a * dy)
    return (*out, converged)


def relax_qp_elastic(Q, q, G, h, penalty, x, t, s1, s2, z1, z2, solver_tol=1e-3, target_kappa=1e-3,
                     consistent=False, A=None, b=None, y=None):
    """qpax/elastic_qp.py:319-359."""
    B = q.shape[0]
    pen = np.broadcast_to(np.asarray(penalty, q.dtype), (B,)).copy()
    if A is None:
        def step(Q, q, G, h, pen, *st):
            return pdip_newton_step_elastic(Q, q, G, h, pen, *st, solver_tol, target_kappa, consistent)
        st, conv, iters = _masked_loop(step, [x, t, s1, s2, z1, z2], [Q, q, G, h, pen])
    else:
        def step(Q, q, G, h, pen, A, b, *st):
            return pdip_newton_step_elastic(Q, q, G, h, pen, *st[:6], solver_tol, target_kappa, consistent,
                                            A=A, b=b, y=st[6])
        st, conv, iters = _masked_loop(step, [x, t, s1, s2, z1, z2, y], [Q, q, G, h, pen, A, b])
    return (*st, conv, iters)


def optnet_derivatives_elastic(dz, dlam, z, lam, exact_dG=False):
    """qpax/elastic_qp.py:366-373 (exact_dG: see optnet_derivatives)."""
    outer = lambda u, v: u[..., :, None] * v[..., None, :]  # noqa: E731
    dl_dQ = 0.5 * (outer(dz, z) + outer(z, dz))
    if exact_dG:
        dl_dG = lam[..., :, None] * outer(dlam, z) + outer(lam, dz)
    else:
        dl_dG = lam[..., :, None] * (outer(dlam, z) + outer(lam, dz))
    return dl_dQ, dz, dl_dG, -lam * dlam


def diff_qp_elastic(Q, q, G, h, x, t, s1, s2, lam1, lam2, dl_dz, exact_dG=False, A=None, y=None):
    """qpax/elastic_qp.py:381-417 -- cotangents of (Q, q, G, h); none for the penalty.  With
    equality rows (extension) also those of (A, b), formed as in optnet_derivatives."""
    nz = q.shape[-1]
    ns = h.shape[-1]
    zns = np.zeros_like(h)
    r7 = None if A is None else np.zeros((q.shape[0], A.shape[-2]), q.dtype)
    dx, dt, ds1, ds2, dlam1, dlam2, _, dnu = solve_elastic_kkt_affine(
        s1, lam1, s2, lam2, Q, G, -dl_dz, zns, zns, zns, zns, zns, A=A, r7=r7)
    dz = np.concatenate((dx, dt), -1)
    z = np.concatenate((x, t), -1)
    lam = np.concatenate((lam1, lam2), -1)
    dlam = np.concatenate((dlam1, dlam2), -1) / lam
    dl_dQ, dl_dq, dl_dG, dl_dh = optnet_derivatives_elastic(dz, dlam, z, lam, exact_dG)
    out = (dl_dQ[:, :nz, :nz], dl_dq[:, :nz], dl_dG[:, ns:, :nz], dl_dh[:, ns:])
    if A is None:
        return out
    outer = lambda u, v: u[..., :, None] * v[..., None, :]  # noqa: E731
    return (*out, outer(dnu, x) + outer(y, dx), -dnu)


def solve_qp_elastic_primal_fwd(Q, q, G, h, penalty, solver_tol=1e-5, target_kappa=1e-3, consistent=False,
                                A=None, b=None):
    """qpax/elastic_qp.py:433-456."""
    Q = np.asarray(Q)
    q = np.asarray(q)
    B, n = q.shape
    Ga = np.asarray(G).reshape(B, -1, n)
    ha = np.asarray(h).reshape(B, -1)
    Ab, bb = _eq_batch(A, b, B, n)
    out = solve_qp_elastic(Q, q, Ga, ha, penalty, solver_tol=solver_tol, A=Ab, b=bb)
    st = out[:-2]
    c1, i1 = out[-2], out[-1]
    y = st[6] if Ab is not None else None
    rel = relax_qp_elastic(Q, q, Ga, ha, penalty, *st[:6], solver_tol=solver_tol, target_kappa=target_kappa,
                           consistent=consistent, A=Ab, b=bb, y=y)
    return st[0], rel[:-2], (c1, i1, rel[-2], rel[-1])


def solve_qp_elastic_primal_bwd(Q, q, G, h, relaxed, dl_dx, exact_dG=False, A=None):
    """qpax/elastic_qp.py:459-466."""
    Q = np.asarray(Q)
    q = np.asarray(q)
    B, n = q.shape
    Ga = np.asarray(G).reshape(B, -1, n)
    ha = np.asarray(h).reshape(B, -1)
    Ab, _ = _eq_batch(A, np.zeros((B, 0)) if A is None else np.zeros((B, np.asarray(A).reshape(B, -1, n).shape[1])), B, n)
    xr, tr, s1r, s2r, z1r, z2r = relaxed[:6]
    yr = relaxed[6] if Ab is not None else None
    return diff_qp_elastic(Q, q, Ga, ha, xr, tr, s1r, s2r, z1r, z2r, dl_dx, exact_dG, A=Ab, y=yr)


def expanded_elastic_problem(Q, q, G, h, penalty, A=None):
    """test/test_elastic_qp.py:39-55 -- the elastic QP written as a standard QP in (x, t); with
    equality rows A (extension) the expanded equality matrix is [A 0]."""
    B, nx = q.shape
    ns = h.shape[-1]
    pen = np.broadcast_to(np.asarray(penalty, q.dtype), (B,))
    Q2 = np.zeros((B, nx + ns, nx + ns), q.dtype)
    Q2[:, :nx, :nx] = Q
    q2 = np.concatenate((q, pen[:, None] * np.ones((B, ns), q.dtype)), -1)
    eye = np.broadcast_to(np.eye(ns, dtype=q.dtype), (B, ns, ns))
    G2 = np.concatenate((np.concatenate((np.zeros((B, ns, nx), q.dtype), -eye), -1),
                         np.concatenate((G, -eye), -1)), 1)
    h2 = np.concatenate((np.zeros((B, ns), q.dtype), h), -1)
    if A is None:
        return Q2, q2, np.zeros((B, 0, nx + ns), q.dtype), np.zeros((B, 0), q.dtype), G2, h2
    A2 = np.concatenate((A, np.zeros((B, A.shape[1], ns), q.dtype)), -1)
    return Q2, q2, A2, None, G2, h2
