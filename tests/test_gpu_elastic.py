"""GPU parity of the elastic QP (qpax/elastic_qp.py) against the float64 oracle, the reference
fixture, and the equality-row extension against the expanded standard-form problem."""

import os

import numpy as np
import pytest

from oracle import qp_oracle as O
from tests import problems as P
from tests._util import rel_err, to_np

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _fixture():
    d = np.load(os.path.join(GOLD, "problem_data_pos.npz"))
    return d, d["Q"][None], d["q"][None], d["G"][None], d["h"][None], float(d["penalty"])


def _t(v, dt=np.float64):
    return torch.tensor(np.asarray(v, dt), device="cuda")


@pytest.mark.parametrize("mode", ["fused", "plain"])
def test_elastic_fixture(mode):
    import qpax_amd

    d, Q, q, G, h, pen = _fixture()
    print(f"[case] elastic fixture mode={mode}", flush=True)
    out = qpax_amd.solve_qp_elastic(_t(Q), _t(q), _t(G), _t(h), pen, solver_tol=1e-10, mode=mode)
    torch.cuda.synchronize()
    assert int(out[6][0]) == 1
    for a, k in zip(out[:6], ("x", "t", "s1", "s2", "z1", "z2")):
        assert np.linalg.norm(to_np(a)[0] - d[k]) < 1e-6, k


def _elastic_batch(B, n, mi, seed, dt):
    Q, q, G, h, shifted = P.elastic_infeasible(B, n, mi, frac=0.3, seed=seed)
    return [np.asarray(v, dt) for v in (Q, q, G, h)]


CASES = [("small-f64", 32, 12, 30, np.float64, 1e-8), ("small-f32", 32, 12, 30, np.float32, 1e-4),
         ("c5-shape-mini", 4, 64, 128, np.float32, 1e-4), ("wide-f64", 4, 100, 200, np.float64, 1e-8),
         ("n70-f64", 4, 70, 90, np.float64, 1e-8)]


@pytest.mark.parametrize("name,B,n,mi,dt,xtol", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("mode", ["fused", "plain"])
def test_elastic_solve_matches_oracle(name, B, n, mi, dt, xtol, mode):
    import qpax_amd

    Q, q, G, h = _elastic_batch(B, n, mi, 3, dt)
    pen = 14.0
    print(f"[case] elastic {name} mode={mode}", flush=True)
    out = qpax_amd.solve_qp_elastic(_t(Q, dt), _t(q, dt), _t(G, dt), _t(h, dt), pen, solver_tol=1e-6, mode=mode)
    torch.cuda.synchronize()
    ref = O.solve_qp_elastic(*[v.astype(np.float64) for v in (Q, q, G, h)], pen, solver_tol=1e-6)
    ok = ref[6] == 1
    assert ok.mean() > 0.8
    np.testing.assert_array_equal(to_np(out[6]), ref[6])
    assert np.abs(to_np(out[7]) - ref[7]).max() <= 1
    assert rel_err(to_np(out[0])[ok], ref[0][ok]).max() < xtol
    assert rel_err(to_np(out[1])[ok], ref[1][ok]).max() < max(xtol, 1e-6) * 100


@pytest.mark.parametrize("relax", ["reference", "consistent"])
@pytest.mark.parametrize("mode", ["fused", "plain"])
def test_elastic_grads_match_oracle(relax, mode):
    import qpax_amd

    Q, q, G, h = _elastic_batch(16, 12, 30, 4, np.float64)
    pen = 14.0
    print(f"[case] elastic grads relax={relax} mode={mode}", flush=True)
    ts = [torch.tensor(v, device="cuda", requires_grad=True) for v in (Q, q, G, h)]
    x = qpax_amd.solve_qp_elastic_primal(*ts, pen, solver_tol=1e-6, target_kappa=1e-3, mode=mode, relax=relax)
    ((x - 1.0) ** 2).sum().backward()
    torch.cuda.synchronize()
    xo, rel, info = O.solve_qp_elastic_primal_fwd(Q, q, G, h, pen, solver_tol=1e-6, target_kappa=1e-3,
                                                  consistent=(relax == "consistent"))
    go = O.solve_qp_elastic_primal_bwd(Q, q, G, h, rel, 2 * (xo - 1))
    ok = info[0] == 1
    assert ok.mean() > 0.8
    # the reference's relaxation (flipped signs) amplifies rounding by ~1.98 per step over its 30
    # steps, so the two implementations' rounding differences grow to ~1e-6 there
    tol = 1e-4 if relax == "reference" else 1e-6
    for nm, t, g in zip("QqGh", ts, go):
        e = rel_err(to_np(t.grad)[ok], g[ok])
        print(f"d{nm}: {e.max():.2e}")
        assert e.max() < tol


def test_elastic_c5_size_f32():
    """The elastic QP at the n=256, m_eq=32, m_ineq=512 size (20% of problems shifted infeasible),
    float32, forward + backward through the elastic primal on a small batch, against the float64
    oracle (matrix-core G^T D G in the workgroup kernels)."""
    import qpax_amd

    Q, q, A, b, G, h, shifted = P.elastic_infeasible(6, 256, 512, frac=0.34, seed=5, me=32)
    pen = 14.0
    print("[case] elastic c5 size B=6", flush=True)
    data = [np.asarray(v, np.float32) for v in (Q, q, A, b, G, h)]
    ts = [torch.tensor(v, device="cuda", requires_grad=True) for v in data]
    x = qpax_amd.solve_qp_elastic_primal(ts[0], ts[1], ts[4], ts[5], pen, A=ts[2], b=ts[3], solver_tol=1e-5,
                                         target_kappa=1e-3)
    ((x - 1.0) ** 2).sum().backward()
    torch.cuda.synchronize()
    d64 = [v.astype(np.float64) for v in data]
    xo, rel, info = O.solve_qp_elastic_primal_fwd(d64[0], d64[1], d64[4], d64[5], pen, solver_tol=1e-5,
                                                  target_kappa=1e-3, A=d64[2], b=d64[3])
    go = O.solve_qp_elastic_primal_bwd(d64[0], d64[1], d64[4], d64[5], rel, 2 * (xo - 1), A=d64[2])
    ok = info[0] == 1
    assert ok.mean() > 0.6
    ex = rel_err(to_np(x).astype(np.float64)[ok], xo[ok])
    print(f"x err max {ex.max():.2e}")
    assert ex.max() < 1e-4
    go = (go[0], go[1], go[4], go[5], go[2], go[3])  # oracle order (Q, q, G, h, A, b)
    for nm, t, g in zip("Q q A b G h".split(), ts, go):
        e = rel_err(to_np(t.grad).astype(np.float64)[ok], g[ok])
        print(f"d{nm}: {e.max():.2e}")
        assert e.max() < 1e-3


@pytest.mark.parametrize("n,me,mi", [(200, 12, 260), (272, 16, 300)], ids=["n200-odd-blocks", "n272-max"])
@pytest.mark.parametrize("mode", ["fused", "plain"])
def test_elastic_tiled_sizes_f64(n, me, mi, mode):
    """Elastic QPs with equality rows at sizes whose blocked Cholesky has an odd number of 16-column
    block columns: n = 200 (13, the last one 8 wide) and n = 272 (17, the largest tiled size), with
    W = L^-1 A^T formed inside the factorisation, forward + backward in float64 against the oracle.
    The relaxation runs with consistent signs: the reference's amplifies rounding differences
    (about 1.98x per step over 30 steps), which would hide the factorisation's accuracy."""
    import qpax_amd

    Q, q, A, b, G, h, shifted = P.elastic_infeasible(2, n, mi, frac=0.5, seed=14, me=me)
    pen = 14.0
    print(f"[case] elastic tiled n={n} me={me} mi={mi} mode={mode}", flush=True)
    ts = [torch.tensor(v, device="cuda", requires_grad=True) for v in (Q, q, A, b, G, h)]
    x = qpax_amd.solve_qp_elastic_primal(ts[0], ts[1], ts[4], ts[5], pen, A=ts[2], b=ts[3], solver_tol=1e-7,
                                         target_kappa=1e-3, mode=mode, relax="consistent")
    ((x - 1.0) ** 2).sum().backward()
    torch.cuda.synchronize()
    xo, rel, info = O.solve_qp_elastic_primal_fwd(Q, q, G, h, pen, solver_tol=1e-7, target_kappa=1e-3,
                                                  consistent=True, A=A, b=b)
    go = O.solve_qp_elastic_primal_bwd(Q, q, G, h, rel, 2 * (xo - 1), A=A)
    ok = info[0] == 1
    assert ok.all()
    ex = rel_err(to_np(x)[ok], xo[ok])
    print(f"x err max {ex.max():.2e}")
    assert ex.max() < 1e-9  # measured 2.5e-11 (n = 200), 8e-13 (n = 272)
    go = (go[0], go[1], go[4], go[5], go[2], go[3])  # oracle order (Q, q, G, h, A, b)
    for nm, t, g in zip("Q q A b G h".split(), ts, go):
        e = rel_err(to_np(t.grad)[ok], g[ok])
        print(f"d{nm}: {e.max():.2e}")
        assert e.max() < 1e-8  # measured at most 6.8e-11


def test_elastic_equality_extension():
    import qpax_amd

    Q, q, A, b, G, h = P.random_feasible(8, 10, 2, 20, seed=9)
    h = h - 2.0 * np.abs(np.random.default_rng(9).standard_normal(h.shape)) * (np.arange(8) % 2)[:, None]
    pen = 5.0
    out = qpax_amd.solve_qp_elastic(_t(Q), _t(q), _t(G), _t(h), pen, solver_tol=1e-9, A=_t(A), b=_t(b))
    torch.cuda.synchronize()
    ref = O.solve_qp_elastic(Q, q, G, h, pen, solver_tol=1e-9, A=A, b=b)
    ok = ref[-2] == 1
    assert ok.mean() > 0.8
    for a, r in zip(out[:7], ref[:7]):
        assert rel_err(to_np(a)[ok], r[ok]).max() < 1e-7
