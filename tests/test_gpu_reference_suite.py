"""The reference's own test suite (test/*.py of qpax), run through qpax_amd on the GPU with the
reference's assertions and tolerances.  Problems are drawn exactly as the reference draws them
(legacy numpy seeding, test/misc_test_utils.py:5-29) or taken from its fixtures."""

import os

import numpy as np
import pytest

from tests import problems as P
from tests._util import to_np

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _t(v, dt=np.float64):
    return torch.tensor(np.asarray(v, dt), device="cuda")


# ------------------------------ test/test_constraint_sizes.py ----------------------------------
M3 = np.array([[1.0, 2.0, 0.0], [-8.0, 3.0, 2.0], [0.0, 1.0, 1.0]])


def test_no_constraints():
    import qpax_amd

    Pm = M3.T @ M3
    q = np.array([3.0, 2.0, 3.0]) @ M3
    x, s, z, y, c, it = qpax_amd.solve_qp(_t(Pm), _t(q), _t(np.zeros((0, 3))), _t(np.zeros(0)),
                                          _t(np.zeros((0, 3))), _t(np.zeros(0)))
    assert int(c) == 1 and int(it) == 0
    assert np.abs(Pm @ to_np(x) + q).max() < 1e-10


def test_eq_only():
    import qpax_amd

    Pm = M3.T @ M3
    q = np.array([3.0, 2.0, 3.0]) @ M3
    x, s, z, y, c, it = qpax_amd.solve_qp(_t(Pm), _t(q), _t(np.array([1.0, 1.0, 1.0])), _t(np.array([1.0])),
                                          _t(np.zeros((0, 3))), _t(np.zeros(0)))
    assert int(c) == 1
    assert np.abs(to_np(x) - [0.28026906, -1.55156951, 2.27130045]).max() < 1e-5


def test_qpsut03():
    import qpax_amd

    Pm = np.array([[122.0, 59, 39, 9], [59, 95, 48, 24], [39, 48, 26, 19], [9, 24, 19, 90]])
    q = np.array([66.0, 93, 47, 11])
    G = np.vstack([-np.eye(4), np.eye(4)])
    h = np.array([np.inf, 0.4, np.inf, 1.0, np.inf, np.inf, 0.5, 1.0])
    x, s, z, y, c, it = qpax_amd.solve_qp(_t(Pm), _t(q), _t(np.zeros((0, 4))), _t(np.zeros(0)), _t(G), _t(h))
    assert int(c) == 1
    assert np.abs(to_np(x) - [0.18143455, 0.00843864, -2.35442995, 0.35443034]).max() < 1e-4


def test_maros_meszaros():
    import qpax_amd

    Pm = np.array([[8.0, 2.0], [2.0, 10.0]])
    q = np.array([1.5, -2.0])
    G = np.array([[-1.0, 2.0], [-2.0, -1.0], [-1.0, 0.0], [0.0, -1.0], [1.0, 0.0], [0.0, 1.0]])
    h = np.array([6.0, -2.0, 0.0, 0.0, 20.0, np.inf])
    x, s, z, y, c, it = qpax_amd.solve_qp(_t(Pm), _t(q), _t(np.zeros((0, 2))), _t(np.zeros(0)), _t(G), _t(h))
    assert int(c) == 1
    assert np.abs(to_np(x) - [0.7625, 0.475]).max() < 1e-4


def test_qptest():
    import qpax_amd

    Qm = np.array([[8.0, 2.0], [2.0, 10.0]])
    q = np.array([1.5, -2.0])
    G = np.vstack([np.array([[-1.0, 2.0], [-2.0, -1.0]]), np.eye(2), -np.eye(2)])
    h = np.concatenate([[6.0, -2.0], [20.0, np.inf], [0.0, 0.0]])
    x, s, z, y, c, it = qpax_amd.solve_qp(_t(Qm), _t(q), _t(np.zeros((0, 2))), _t(np.zeros(0)), _t(G), _t(h),
                                          solver_tol=1e-6)
    x, s, z = to_np(x), to_np(s), to_np(z)
    res = np.concatenate([Qm @ x + q + G.T @ z, s * z, G @ x + s - h])
    res = np.where(np.isnan(res), 0, res)
    assert int(c) == 1 and np.abs(res).max() < 1e-5


# ------------------------------ test/test_qp_solver_{double,single}.py -------------------------
def _reference_batch(ny, count=1000, seed=1):
    rs = np.random.RandomState(seed)
    probs = [P.planted_kkt_legacy(15, 10, ny, rs) for _ in range(count)]
    return [np.stack([p[i] for p in probs]) for i in range(6)]


def _kkt(Q, q, A, b, G, h, x, s, z, y):
    r1 = np.einsum("bij,bj->bi", Q, x) + q + np.einsum("bji,bj->bi", A, y) + np.einsum("bji,bj->bi", G, z)
    r3 = np.einsum("bij,bj->bi", G, x) + s - h
    r4 = np.einsum("bij,bj->bi", A, x) - b
    return np.abs(r1).max(1), np.abs(s * z).max(1), np.abs(r3).max(1), (np.abs(r4).max(1) if b.shape[1] else 0 * r3[:, 0])


@pytest.mark.parametrize("dt,tol,maxit", [(np.float64, 1e-6, 20), (np.float32, 1e-2, 10)], ids=["double", "single"])
@pytest.mark.parametrize("ny", [3, 0])
@pytest.mark.parametrize("mode", ["fused", "plain"])
def test_qp_solver(dt, tol, maxit, ny, mode):
    import qpax_amd

    data = _reference_batch(ny, seed=1)
    print(f"[case] reference solver suite dt={dt.__name__} ny={ny} mode={mode}", flush=True)
    out = qpax_amd.solve_qp(*[_t(v, dt) for v in data], solver_tol=tol, mode=mode)
    torch.cuda.synchronize()
    x, s, z, y, c, it = [to_np(v).astype(np.float64) for v in out]
    assert np.all(c == 1)
    assert it.max() <= maxit
    for r in _kkt(*[v.astype(dt).astype(np.float64) for v in data], x, s, z, y):
        assert r.max() <= tol


# ------------------------------ test/test_qp_derivatives.py ------------------------------------
def test_derivs_reference_problem():
    import qpax_amd
    from oracle import qp_oracle as O

    prob = P.planted_kkt_legacy(15, 10, 3, np.random.RandomState(3))
    data = [np.asarray(v)[None] for v in prob[:6]]
    ts = [torch.tensor(v, device="cuda", requires_grad=True) for v in data]
    x = qpax_amd.solve_qp_primal(*ts, target_kappa=1e-3)
    ((x - 1.0) ** 2).sum().backward()
    grads = [to_np(t.grad) for t in ts]

    def loss(batch):
        xx = O.solve_qp(*batch, solver_tol=1e-5)[0]
        return ((xx - 1) ** 2).sum(-1)

    f0 = loss(data)[0]
    for ai, (arr, g) in enumerate(zip(data, grads)):
        pert = []
        for e in range(arr[0].size):
            p = [a.copy() for a in data]
            p[ai].reshape(1, -1)[0, e] += 1e-3
            pert.append(p)
        batch = [np.concatenate([p[k] for p in pert]) for k in range(6)]
        fd = ((loss(batch) - f0) / 1e-3).reshape(arr.shape)
        assert np.linalg.norm(g - fd) < 0.2 * np.linalg.norm(fd)


# ------------------------------ test/test_elastic_qp.py ----------------------------------------
def test_elastic_fixture_suite():
    import qpax_amd
    from oracle import qp_oracle as O

    d = np.load(os.path.join(GOLD, "problem_data_pos.npz"))
    Q, q, G, h, pen = d["Q"], d["q"], d["G"], d["h"], float(d["penalty"])
    Q2, q2, A2, b2, G2, h2 = [v[0] for v in O.expanded_elastic_problem(Q[None], q[None], G[None], h[None], pen)]
    xb, sb, zb, yb, cb, ib = qpax_amd.solve_qp(_t(Q2), _t(q2), _t(A2), _t(b2), _t(G2), _t(h2), solver_tol=1e-10)
    assert int(cb) == 1
    assert np.linalg.norm(to_np(xb) - np.concatenate((d["x"], d["t"]))) < 1e-6
    assert np.linalg.norm(to_np(sb) - np.concatenate((d["s1"], d["s2"]))) < 1e-6
    assert np.linalg.norm(to_np(zb) - np.concatenate((d["z1"], d["z2"]))) < 1e-6
    out = qpax_amd.solve_qp_elastic(_t(Q), _t(q), _t(G), _t(h), pen, solver_tol=1e-10)
    assert int(out[6]) == 1
    for a, k in zip(out[:6], ("x", "t", "s1", "s2", "z1", "z2")):
        assert np.linalg.norm(to_np(a) - d[k]) < 1e-6
    xp = qpax_amd.solve_qp_elastic_primal(_t(Q), _t(q), _t(G), _t(h), pen, solver_tol=1e-10)
    assert np.linalg.norm(d["x"] - to_np(xp)) < 1e-6
    # grads of the elastic primal == grads of the expanded plain primal (test_elastic_qp.py:113-144)
    x_des = np.random.default_rng(0).standard_normal(5)
    t1 = [torch.tensor(v, device="cuda", requires_grad=True) for v in (Q, q, G, h)]
    xe = qpax_amd.solve_qp_elastic_primal(*t1, pen, solver_tol=1e-6, target_kappa=1e-3)
    ((xe - _t(x_des)) ** 2).sum().backward()
    t2 = [torch.tensor(v, device="cuda", requires_grad=True) for v in (Q, q, G, h)]
    Q2t = torch.zeros((25, 25), dtype=torch.float64, device="cuda")
    Q2t[:5, :5] = t2[0]
    q2t = torch.cat([t2[1], pen * torch.ones(20, dtype=torch.float64, device="cuda")])
    eye = torch.eye(20, dtype=torch.float64, device="cuda")
    G2t = torch.cat([torch.cat([torch.zeros((20, 5), dtype=torch.float64, device="cuda"), -eye], 1),
                     torch.cat([t2[2], -eye], 1)], 0)
    h2t = torch.cat([torch.zeros(20, dtype=torch.float64, device="cuda"), t2[3]])
    xo = qpax_amd.solve_qp_primal(Q2t, q2t, _t(np.zeros((0, 25))), _t(np.zeros(0)), G2t, h2t, solver_tol=1e-6,
                                  target_kappa=1e-3)
    ((xo[:5] - _t(x_des)) ** 2).sum().backward()
    for a, b in zip(t1, t2):
        assert np.linalg.norm(to_np(a.grad) - to_np(b.grad)) < 1e-3
