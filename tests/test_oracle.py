"""Pins the float64 CPU oracle (oracle/qp_oracle.py) before it is trusted as the GPU checker.

Each check is independent of the oracle's own formulas: the reference's fixtures and known
answers, closed-form optima, a library routine the problem reduces to, an independent dense
Newton solve, and finite differences.  The last tests plant a mistake in a copy of the oracle and
show that these checks catch it.
"""

import os

import numpy as np
import pytest

from oracle import qp_oracle as O
from tests import problems as P

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _fixture():
    d = np.load(os.path.join(GOLD, "problem_data_pos.npz"))
    iv = np.load(os.path.join(GOLD, "initial_values_pos.npz"))
    Q, q, G, h = d["Q"][None], d["q"][None], d["G"][None], d["h"][None]
    return d, iv, Q, q, G, h, float(d["penalty"])


# --------------------------- reference fixtures (test/test_elastic_qp.py) -----------------------
def test_expanded_problem_matches_fixture_arrays():
    d, iv, Q, q, G, h, pen = _fixture()
    Q2, q2, A2, b2, G2, h2 = O.expanded_elastic_problem(Q, q, G, h, pen)
    np.testing.assert_array_equal(Q2[0], d["Q2"])
    np.testing.assert_array_equal(q2[0], d["q2"])
    np.testing.assert_array_equal(G2[0], d["G2"])
    np.testing.assert_array_equal(h2[0], d["h2"])


def test_plain_solve_matches_fixture():  # test_elastic_qp.py:64-73
    d, iv, Q, q, G, h, pen = _fixture()
    x, s, z, y, c, it = O.solve_qp(*O.expanded_elastic_problem(Q, q, G, h, pen), solver_tol=1e-10)
    assert c[0] == 1
    assert np.linalg.norm(x[0] - np.concatenate([d["x"], d["t"]])) < 1e-6
    assert np.linalg.norm(s[0] - np.concatenate([d["s1"], d["s2"]])) < 1e-6
    assert np.linalg.norm(z[0] - np.concatenate([d["z1"], d["z2"]])) < 1e-6


def test_plain_initialize_matches_fixture():  # test_elastic_qp.py:79-82
    d, iv, Q, q, G, h, pen = _fixture()
    Q2, q2, A2, b2, G2, h2 = O.expanded_elastic_problem(Q, q, G, h, pen)
    x, s, z, y = O.initialize(Q2, q2, A2, b2, G2, h2)
    assert np.linalg.norm(x[0] - np.concatenate([iv["x"], iv["t"]])) < 1e-6
    assert np.linalg.norm(s[0] - np.concatenate([iv["s1"], iv["s2"]])) < 1e-6
    assert np.linalg.norm(z[0] - np.concatenate([iv["z1"], iv["z2"]])) < 1e-6


def test_elastic_initialize_and_solve_match_fixture():  # test_elastic_qp.py:89-111
    d, iv, Q, q, G, h, pen = _fixture()
    init = O.initialize_elastic(Q, q, G, h, pen)
    for a, k in zip(init, ("x", "t", "s1", "s2", "z1", "z2")):
        assert np.linalg.norm(a[0] - iv[k]) < 1e-6
    out = O.solve_qp_elastic(Q, q, G, h, pen, solver_tol=1e-10)
    assert out[6][0] == 1
    for a, k in zip(out[:6], ("x", "t", "s1", "s2", "z1", "z2")):
        assert np.linalg.norm(a[0] - d[k]) < 1e-6
    x, _, _ = O.solve_qp_elastic_primal_fwd(Q, q, G, h, pen, solver_tol=1e-10)
    assert np.linalg.norm(x[0] - d["x"]) < 1e-6


@pytest.mark.parametrize("consistent", [False, True])
def test_elastic_grads_equal_expanded_plain_grads(consistent):  # test_elastic_qp.py:113-144
    d, iv, Q, q, G, h, pen = _fixture()
    x_des = np.random.default_rng(0).standard_normal(5)
    xe, rel, _ = O.solve_qp_elastic_primal_fwd(Q, q, G, h, pen, solver_tol=1e-6, target_kappa=1e-3,
                                               consistent=consistent)
    ge = O.solve_qp_elastic_primal_bwd(Q, q, G, h, rel, 2 * (xe - x_des))
    Q2, q2, A2, b2, G2, h2 = O.expanded_elastic_problem(Q, q, G, h, pen)
    xp, relp, _ = O.solve_qp_primal_fwd(Q2, q2, A2, b2, G2, h2, solver_tol=1e-6, target_kappa=1e-3)
    v = np.zeros_like(xp)
    v[:, :5] = 2 * (xp[:, :5] - x_des)
    gp = O.solve_qp_primal_bwd(Q2, q2, A2, b2, G2, h2, relp, v)
    ref = (gp[0][:, :5, :5], gp[1][:, :5], gp[4][:, 20:, :5], gp[5][:, 20:])
    for a, b in zip(ge, ref):
        assert np.linalg.norm(a - b) < 1e-3


# --------------------------- known answers (test/test_constraint_sizes.py) ----------------------
M3 = np.array([[1.0, 2.0, 0.0], [-8.0, 3.0, 2.0], [0.0, 1.0, 1.0]])


def test_no_constraints():  # :11-27
    Pm = M3.T @ M3
    q = np.array([3.0, 2.0, 3.0]) @ M3
    x, s, z, y, c, it = O.solve_qp(Pm[None], q[None], np.zeros((1, 0, 3)), np.zeros((1, 0)),
                                   np.zeros((1, 0, 3)), np.zeros((1, 0)))
    assert c[0] == 1 and it[0] == 0
    assert np.abs(Pm @ x[0] + q).max() < 1e-10


def test_eq_only():  # :30-61
    Pm = M3.T @ M3
    q = np.array([3.0, 2.0, 3.0]) @ M3
    A = np.ones((1, 1, 3))
    x, s, z, y, c, it = O.solve_qp(Pm[None], q[None], A, np.ones((1, 1)), np.zeros((1, 0, 3)), np.zeros((1, 0)))
    assert c[0] == 1
    assert np.abs(x[0] - [0.28026906, -1.55156951, 2.27130045]).max() < 1e-5
    K = np.block([[Pm, A[0].T], [A[0], np.zeros((1, 1))]])
    sol = np.linalg.solve(K, np.concatenate([-q, [1.0]]))
    assert np.abs(x[0] - sol[:3]).max() < 1e-10 and np.abs(y[0] - sol[3:]).max() < 1e-10


def test_qpsut03_inf_bounds():  # :64-100
    Pm = np.array([[122.0, 59, 39, 9], [59, 95, 48, 24], [39, 48, 26, 19], [9, 24, 19, 90]])
    q = np.array([66.0, 93, 47, 11])
    G = np.vstack([-np.eye(4), np.eye(4)])
    h = np.array([np.inf, 0.4, np.inf, 1.0, np.inf, np.inf, 0.5, 1.0])
    x, s, z, y, c, it = O.solve_qp(Pm[None], q[None], np.zeros((1, 0, 4)), np.zeros((1, 0)), G[None], h[None])
    assert c[0] == 1
    assert np.abs(x[0] - [0.18143455, 0.00843864, -2.35442995, 0.35443034]).max() < 1e-4
    assert np.all(z[0][np.isinf(h)] == 0) and np.all(np.isinf(s[0][np.isinf(h)]))


def test_maros_meszaros_inf_bound():  # :103-132
    Pm = np.array([[8.0, 2.0], [2.0, 10.0]])
    q = np.array([1.5, -2.0])
    G = np.array([[-1.0, 2.0], [-2.0, -1.0], [-1.0, 0.0], [0.0, -1.0], [1.0, 0.0], [0.0, 1.0]])
    h = np.array([6.0, -2.0, 0.0, 0.0, 20.0, np.inf])
    x, s, z, y, c, it = O.solve_qp(Pm[None], q[None], np.zeros((1, 0, 2)), np.zeros((1, 0)), G[None], h[None])
    assert c[0] == 1
    assert np.abs(x[0] - [0.7625, 0.475]).max() < 1e-4


def test_qptest_kkt():  # :135-163
    Qm = np.array([[8.0, 2.0], [2.0, 10.0]])
    q = np.array([1.5, -2.0])
    G = np.vstack([np.array([[-1.0, 2.0], [-2.0, -1.0]]), np.eye(2), -np.eye(2)])
    h = np.concatenate([[6.0, -2.0], [20.0, np.inf], [0.0, 0.0]])
    x, s, z, y, c, it = O.solve_qp(Qm[None], q[None], np.zeros((1, 0, 2)), np.zeros((1, 0)), G[None], h[None],
                                   solver_tol=1e-6)
    x, s, z = x[0], s[0], z[0]
    res = np.concatenate([Qm @ x + q + G.T @ z, s * z, G @ x + s - h])
    res = np.where(np.isnan(res), 0, res)
    assert c[0] == 1 and np.abs(res).max() < 1e-5


# --------------------------- closed forms and library routines ----------------------------------
@pytest.mark.parametrize("ny", [3, 0])
def test_planted_kkt_point(ny):
    """generate_random_qp (test/misc_test_utils.py:5-29) plants a strictly complementary KKT point;
    tol 1e-6, <= 20 iterations and the KKT check as in test/test_qp_solver_double.py."""
    Q, q, A, b, G, h, xt, st, zt, yt = P.planted_kkt(200, 15, 10, ny, seed=11)
    x, s, z, y, c, it = O.solve_qp(Q, q, A, b, G, h, solver_tol=1e-6)
    assert np.all(c == 1) and it.max() <= 20
    r1 = np.einsum("bij,bj->bi", Q, x) + q + np.einsum("bji,bj->bi", A, y) + np.einsum("bji,bj->bi", G, z)
    r3 = np.einsum("bij,bj->bi", G, x) + s - h
    assert np.abs(r1).max() <= 1e-6 and np.abs(s * z).max() <= 1e-6 and np.abs(r3).max() <= 1e-6
    # Q = M^T M is a.s. positive definite, so the optimum is unique: x equals the planted point
    assert np.abs(x - xt).max() < 1e-4


def test_nnls_matches_scipy():
    from scipy.optimize import nnls

    rng = np.random.default_rng(1)
    F = rng.standard_normal((64, 10, 5))
    g = rng.standard_normal((64, 10))
    Q = np.swapaxes(F, 1, 2) @ F
    qv = -np.einsum("bji,bj->bi", F, g)
    G = np.broadcast_to(-np.eye(5), (64, 5, 5)).copy()
    x, s, z, y, c, it = O.solve_qp(Q, qv, np.zeros((64, 0, 5)), np.zeros((64, 0)), G, np.zeros((64, 5)),
                                   solver_tol=1e-10)
    assert np.all(c == 1)
    ref = np.stack([nnls(F[i], g[i])[0] for i in range(64)])
    assert np.abs(x - ref).max() < 1e-7


# --------------------------- relaxation and backward --------------------------------------------
def _dense_relaxed_newton(Q, q, A, b, G, h, kappa, x, s, z, y, iters=60):
    """Independent check: undamped-to-damped Newton on the full (unreduced) relaxed KKT system
    with numpy's dense LU, for one problem."""
    n, me, mi = len(q), len(b), len(h)
    for _ in range(iters):
        r = np.concatenate([Q @ x + q + A.T @ y + G.T @ z, s * z - kappa, G @ x + s - h, A @ x - b])
        K = np.zeros((n + 2 * mi + me,) * 2)
        K[:n, :n] = Q
        K[:n, n + mi:n + 2 * mi] = G.T
        K[:n, n + 2 * mi:] = A.T
        K[n:n + mi, n:n + mi] = np.diag(z)
        K[n:n + mi, n + mi:n + 2 * mi] = np.diag(s)
        K[n + mi:n + 2 * mi, :n] = G
        K[n + mi:n + 2 * mi, n:n + mi] = np.eye(mi)
        K[n + 2 * mi:, :n] = A
        d = np.linalg.solve(K, -r)
        dx, ds, dz, dy = d[:n], d[n:n + mi], d[n + mi:n + 2 * mi], d[n + 2 * mi:]
        a = 1.0
        for v, dv in ((s, ds), (z, dz)):
            neg = dv < 0
            if neg.any():
                a = min(a, 0.99 * np.min(-v[neg] / dv[neg]))
        x, s, z, y = x + a * dx, s + a * ds, z + a * dz, y + a * dy
    return x, s, z, y


def test_relax_reaches_central_path_point():
    Q, q, A, b, G, h, *_ = P.planted_kkt(8, 6, 8, 2, seed=4)
    kappa = 1e-3
    xo, (xr, sr, zr, yr), info = O.solve_qp_primal_fwd(Q, q, A, b, G, h, solver_tol=1e-10, target_kappa=kappa)
    assert np.all(info[2] == 1)
    assert np.abs(sr * zr - kappa).max() < 1e-9
    for i in range(8):
        xd, sd, zd, yd = _dense_relaxed_newton(Q[i], q[i], A[i], b[i], G[i], h[i], kappa, xo[i], np.ones(8),
                                               np.ones(8), np.zeros(2))
        assert np.abs(xd - xr[i]).max() < 1e-8 and np.abs(zd - zr[i]).max() < 1e-8


def _relaxed_x(Q, q, A, b, G, h, kappa):
    _, (xr, _, _, _), info = O.solve_qp_primal_fwd(Q, q, A, b, G, h, solver_tol=1e-12, target_kappa=kappa)
    return xr, info


def test_backward_is_vjp_of_relaxed_map():
    """diff_qp returns the exact VJP of the relaxed solution map x_kappa(theta): central finite
    differences (step 1e-6) of <v, x_kappa> over every input element.  For G this holds for the
    exact_dG form; the reference's dG formula (diff_qp.py:11) scales its second term by an extra
    lam, which the last assertion shows is measurable."""
    # (the damped relaxation does not converge from every planted point within 30 steps -- the
    # reference's behaviour too -- so the problem is one where it does)
    Q, q, A, b, G, h, *_ = P.planted_kkt(1, 4, 6, 1, seed=9)
    kappa = 1e-2
    args = [Q, q, A, b, G, h]
    x0, info = _relaxed_x(*args, kappa)
    assert info[2][0] == 1
    v = np.random.default_rng(2).standard_normal((1, 4))
    _, rel, _ = O.solve_qp_primal_fwd(*args, solver_tol=1e-12, target_kappa=kappa)
    grads = O.solve_qp_primal_bwd(*args, rel, v, exact_dG=True)
    ref_dG = O.solve_qp_primal_bwd(*args, rel, v)[4]
    eps = 1e-6
    for ai, (arr, g) in enumerate(zip(args, grads)):
        flat = arr.reshape(1, -1)
        pert = []
        for e in range(flat.shape[1]):
            for sgn in (1, -1):
                p = [a.copy() for a in args]
                pf = p[ai].reshape(1, -1)
                pf[0, e] += sgn * eps
                pert.append(p)
        batch = [np.concatenate([p[k] for p in pert]) for k in range(6)]
        xs, inf2 = _relaxed_x(*batch, kappa)
        assert np.all(inf2[2] == 1)
        f = xs @ v[0]
        fd = ((f[0::2] - f[1::2]) / (2 * eps)).reshape(arr.shape)
        if ai == 0:  # the solver symmetrises Q: compare the symmetric part
            fd = 0.5 * (fd + np.swapaxes(fd, -1, -2))
        assert np.abs(fd - g).max() < 1e-5 * max(1.0, np.abs(g).max()), "QqAbGh"[ai]
        if ai == 4:
            assert np.abs(fd - ref_dG).max() > 1e-3 * np.abs(fd).max()


def test_backward_fd_unrelaxed_loose():
    """The reference's own gradient test (test/test_qp_derivatives.py) on its own problem
    (np.random.seed(3), generate_random_qp(15, 10, 3)): grads of ||x-1||^2 at target_kappa=1e-3
    vs forward differences (step 1e-3) of the unrelaxed solution, within 20%."""
    prob = P.planted_kkt_legacy(15, 10, 3, np.random.RandomState(3))
    args = [np.asarray(v)[None] for v in prob[:6]]
    x, rel, _ = O.solve_qp_primal_fwd(*args, solver_tol=1e-5, target_kappa=1e-3)
    grads = O.solve_qp_primal_bwd(*args, rel, 2 * (x - 1))

    def loss(batch):
        xx = O.solve_qp(*batch, solver_tol=1e-5)[0]
        return ((xx - 1) ** 2).sum(-1)

    f0 = loss(args)[0]
    for ai, (arr, g) in enumerate(zip(args, grads)):
        nel = arr[0].size
        pert = []
        for e in range(nel):
            p = [a.copy() for a in args]
            p[ai].reshape(1, -1)[0, e] += 1e-3
            pert.append(p)
        batch = [np.concatenate([p[k] for p in pert]) for k in range(6)]
        fd = ((loss(batch) - f0) / 1e-3).reshape(arr.shape)
        assert np.linalg.norm(g - fd) < 0.2 * np.linalg.norm(fd), ("QqAbGh"[ai], np.linalg.norm(g - fd), np.linalg.norm(fd))


# --------------------------- planted mistakes are caught -----------------------------------------
def test_planted_mistake_in_step_rule_is_caught(monkeypatch):
    """Dropping the 0.99 fraction-to-boundary factor of the predictor-corrector step changes the
    trajectory; the reference fixture then fails at its own tolerance."""
    orig = O.ort_linesearch
    calls = {"n": 0}

    def wrong_ort(x, dx):
        calls["n"] += 1
        return np.minimum(1.0, 1.1 * orig(x, dx))  # overshoot the boundary slightly

    monkeypatch.setattr(O, "ort_linesearch", wrong_ort)
    d, iv, Q, q, G, h, pen = _fixture()
    with np.errstate(all="ignore"):
        x, s, z, y, c, it = O.solve_qp(*O.expanded_elastic_problem(Q, q, G, h, pen), solver_tol=1e-10)
    bad = (c[0] != 1) or np.linalg.norm(x[0] - np.concatenate([d["x"], d["t"]])) >= 1e-6 or \
        not np.all(np.isfinite(x))
    assert calls["n"] > 0 and bad


def test_planted_mistake_in_backward_sign_is_caught(monkeypatch):
    """A sign error in optnet_derivatives' dh is caught by the finite-difference check."""
    orig = O.optnet_derivatives

    def wrong(dz, dlam, dnu, z, lam, nu, exact_dG=False):
        out = list(orig(dz, dlam, dnu, z, lam, nu, exact_dG))
        out[5] = -out[5]
        return tuple(out)

    monkeypatch.setattr(O, "optnet_derivatives", wrong)
    with pytest.raises(AssertionError):
        test_backward_is_vjp_of_relaxed_map()


# --------------------------- elastic equality extension ------------------------------------------
def _elastic_eq_problem(B=6, n=8, me=2, mi=12, seed=21, shift=True):
    Q, q, A, b, G, h = P.random_feasible(B, n, me, mi, seed=seed)
    if shift:  # make some problems infeasible so the slacks t are active
        h = h - 2.0 * np.abs(np.random.default_rng(seed).standard_normal(h.shape)) * (np.arange(B) % 2)[:, None]
    return Q, q, A, b, G, h


def test_elastic_equality_extension_matches_expanded_problem():
    Q, q, A, b, G, h = _elastic_eq_problem()
    pen = 5.0
    x, t, s1, s2, z1, z2, y, c, it = O.solve_qp_elastic(Q, q, G, h, pen, solver_tol=1e-10, A=A, b=b)
    Q2, q2, A2, _, G2, h2 = O.expanded_elastic_problem(Q, q, G, h, pen, A=A)
    xe, se, ze, ye, ce, ie = O.solve_qp(Q2, q2, A2, b, G2, h2, solver_tol=1e-10)
    ok = (c == 1) & (ce == 1)
    assert ok.mean() > 0.8
    assert np.abs(x - xe[:, :8])[ok].max() < 1e-7
    assert np.abs(t - xe[:, 8:])[ok].max() < 1e-7
    assert np.abs(y - ye)[ok].max() < 1e-6
    assert np.abs(np.concatenate([z1, z2], -1) - ze)[ok].max() < 1e-6


@pytest.mark.parametrize("consistent", [False, True])
def test_elastic_equality_extension_grads_match_expanded(consistent):
    Q, q, A, b, G, h = _elastic_eq_problem(seed=22)
    pen = 5.0
    v = np.random.default_rng(5).standard_normal((Q.shape[0], 8))
    x, rel, info = O.solve_qp_elastic_primal_fwd(Q, q, G, h, pen, solver_tol=1e-8, target_kappa=1e-3,
                                                 consistent=consistent, A=A, b=b)
    ge = O.solve_qp_elastic_primal_bwd(Q, q, G, h, rel, v, A=A)
    Q2, q2, A2, _, G2, h2 = O.expanded_elastic_problem(Q, q, G, h, pen, A=A)
    xp, relp, infop = O.solve_qp_primal_fwd(Q2, q2, A2, b, G2, h2, solver_tol=1e-8, target_kappa=1e-3)
    vv = np.zeros_like(xp)
    vv[:, :8] = v
    gp = O.solve_qp_primal_bwd(Q2, q2, A2, b, G2, h2, relp, vv)
    ok = (info[0] == 1) & (infop[0] == 1) & (infop[2] == 1)
    ref = (gp[0][:, :8, :8], gp[1][:, :8], gp[4][:, 12:, :8], gp[5][:, 12:], gp[2][:, :, :8], gp[3])
    assert ok.mean() > 0.5
    for a, r in zip(ge, ref):
        assert np.abs(a - r)[ok].max() < 1e-3 * max(1.0, np.abs(r[ok]).max())


def test_conditioning_floor_separates_well_and_ill_conditioned_batches():
    """tests._util.conditioning_floor (the float64 bar of the GPU state tests) stays far below 1e-9
    on well-conditioned problems, so there the bar is the fixed 1e-9, and rises above it on
    problems whose KKT systems are ill-conditioned at the stopping point (n=52, m_ineq=256)."""
    from tests._util import conditioning_floor

    for (n, me, mi, seed, tol), lo, hi in (((20, 5, 40, 2, 1e-3), 0.0, 1e-9), ((52, 8, 256, 39, 1e-6), 1e-9, 1e-6)):
        data = list(P.random_feasible(16, n, me, mi, seed=seed))
        ref = O.solve_qp(*data, solver_tol=tol)
        fl = conditioning_floor(lambda *v: O.solve_qp(*v, solver_tol=tol), data, ref, ref[4] == 1)
        worst = max(fl.values())
        assert lo <= worst < hi, (n, fl)


def test_trace_follows_the_solve_and_its_stopping_rule():
    """solve_qp_trace: with solve_qp's step rule its iterate and flags are solve_qp's exactly; each
    problem has `iters` trace rows, the last row's residual norms fall below the tolerance exactly
    for the converged problems (the loop stops on the entry test, pdip.py:100), and every step
    length lies in (0, 0.99].  With the debug loop's rule the step lengths are capped at
    0.99 * 0.99 and the solution is the same QP's."""
    data = list(P.random_feasible(24, 20, 5, 40, seed=2))
    ref = O.solve_qp(*data, solver_tol=1e-6)
    out, tr = O.solve_qp_trace(*data, solver_tol=1e-6)
    for a, b in zip(out, ref):
        np.testing.assert_array_equal(a, b)
    iters = out[5]
    for p in range(24):
        k = iters[p]
        assert np.isfinite(tr[p, :k]).all() and np.isnan(tr[p, k:]).all()
        last = tr[p, k - 1, :4].max()
        assert (last < 1e-6) == (out[4][p] == 1)
        assert ((tr[p, :k, 4] > 0) & (tr[p, :k, 4] <= 0.99)).all()
        assert (tr[p, :k - 1, :4].max(axis=1) >= 1e-6).all()
    dbg, trd = O.solve_qp_trace(*data, solver_tol=1e-6, debug_step=True)
    kd = dbg[5]
    assert all((trd[p, :kd[p], 4] <= 0.99 * 0.99 + 1e-15).all() for p in range(24))
    ok = (dbg[4] == 1) & (ref[4] == 1)
    assert np.abs(dbg[0][ok] - ref[0][ok]).max() < 1e-4
    # no equality rows: |r4| is 0, as the debug print's zeros(1)
    nn = list(P.nnls(4, seed=1))
    _, trn = O.solve_qp_trace(*nn, solver_tol=1e-8)
    assert np.all(trn[:, 0, 3] == 0.0)
