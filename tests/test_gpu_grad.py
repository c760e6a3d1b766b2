"""GPU parity of solve_qp_primal's forward relaxation and backward against the float64 oracle."""

import numpy as np
import pytest

from oracle import qp_oracle as O
from tests import problems as P
from tests._util import rel_err, to_np

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

GCASES = [
    ("c3-mini-f64", lambda: P.random_feasible(32, 50, 10, 100, seed=3), np.float64, 1e-7),
    ("c3-mini-f32", lambda: P.random_feasible(32, 50, 10, 100, seed=3), np.float32, 1e-3),
    ("c2-mini-f32", lambda: P.random_feasible(32, 20, 5, 40, seed=2), np.float32, 1e-3),
    ("planted-f64", lambda: P.planted_kkt(32, 15, 10, 3, seed=3)[:6], np.float64, 1e-7),
    ("mid-f32", lambda: P.random_feasible(16, 100, 20, 250, seed=6), np.float32, 1e-3),
    # n = 70: the relaxation's stopping test (residual < 1e-5) can flip by one iteration between two
    # fp64 eliminations on a few problems, which moves the relaxed point by a step of order 1e-6
    ("odd-f64", lambda: P.random_feasible(16, 70, 3, 140, seed=8), np.float64, 1e-5),
    # box rows (G = [G_dense; I; -I]): the wavefront kernels' diagonal/gather path for the +-e_j rows
    ("box20-f32", lambda: P.box_qp(32, 20, 3, 10, seed=5), np.float32, 1e-3),
    ("box50-f64", lambda: P.box_qp(32, 50, 8, 20, seed=7), np.float64, 1e-7),
    ("nnls-f64", lambda: P.nnls(64, seed=1), np.float64, 1e-7),
]


def _grads(data, dt, mode):
    import qpax_amd

    ts = [torch.tensor(np.asarray(v, dt), device="cuda", requires_grad=True) for v in data]
    x = qpax_amd.solve_qp_primal(*ts, solver_tol=1e-5, target_kappa=1e-3, mode=mode)
    loss = ((x - 1.0) ** 2).sum()
    loss.backward()
    torch.cuda.synchronize()
    return to_np(x), [to_np(t.grad) for t in ts]


@pytest.mark.parametrize("name,gen,dt,gtol", GCASES, ids=[c[0] for c in GCASES])
def test_grads_match_oracle(name, gen, dt, gtol):
    data = [np.asarray(v, dt) for v in gen()]
    print(f"[case] {name} shapes={[d.shape for d in data]}", flush=True)
    x, g = _grads(data, dt, "fused")
    d64 = [v.astype(np.float64) for v in data]
    xo, rel, info = O.solve_qp_primal_fwd(*d64, solver_tol=1e-5, target_kappa=1e-3)
    go = O.solve_qp_primal_bwd(*d64, rel, 2 * (xo - 1))
    ok = (info[0] == 1) & (info[2] == 1)
    assert ok.mean() > 0.9
    for nm, a, b in zip("Q q A b G h".split(), g, go):
        if b.shape[1] == 0:
            continue
        e = rel_err(a[ok], b[ok])
        print(f"{name} d{nm}: max {e.max():.2e} median {np.median(e):.2e}")
        assert e.max() < gtol
        assert np.median(e) < (1e-8 if dt == np.float64 else 1e-4)


def test_grads_plain_agrees_with_fused():
    data = [np.asarray(v, np.float64) for v in P.random_feasible(16, 50, 10, 100, seed=7)]
    xa, ga = _grads(data, np.float64, "fused")
    xb, gb = _grads(data, np.float64, "plain")
    assert rel_err(xa, xb).max() < 1e-9
    for a, b in zip(ga, gb):
        assert rel_err(a, b).max() < 1e-6


def test_exact_dG_matches_oracle():
    import qpax_amd

    data = [np.asarray(v, np.float64) for v in P.random_feasible(16, 50, 10, 100, seed=12)]
    ts = [torch.tensor(v, device="cuda", requires_grad=True) for v in data]
    x = qpax_amd.solve_qp_primal(*ts, solver_tol=1e-6, target_kappa=1e-3, exact_dG=True)
    ((x - 1.0) ** 2).sum().backward()
    torch.cuda.synchronize()
    xo, rel, info = O.solve_qp_primal_fwd(*data, solver_tol=1e-6, target_kappa=1e-3)
    go = O.solve_qp_primal_bwd(*data, rel, 2 * (xo - 1), exact_dG=True)
    ok = (info[0] == 1) & (info[2] == 1)
    # at solver_tol 1e-6 the reduced KKT matrix is conditioned so that a different elimination's
    # rounding moves the relaxed point (and dG with it) by up to ~1e-6 on a few problems
    e = rel_err(to_np(ts[4].grad)[ok], go[4][ok])
    assert e.max() < 1e-6 and np.median(e) < 1e-8
    # and the exact form is told apart from the reference's combination (off by percents)
    gref = O.solve_qp_primal_bwd(*data, rel, 2 * (xo - 1), exact_dG=False)[4]
    assert np.median(rel_err(to_np(ts[4].grad)[ok], gref[ok])) > 1e-3
