"""GPU parity of the forward solve and the relaxation against the float64 oracle."""

import numpy as np
import pytest

from oracle import qp_oracle as O
from qpax_amd import problems as P
from tests._util import rel_err, to_np

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

CASES = [
    # (name, generator, dtype, tol, x_tol)
    ("c3-mini-f64", lambda: P.random_feasible(64, 50, 10, 100, seed=3), np.float64, 1e-5, 1e-9),
    ("c3-mini-f32", lambda: P.random_feasible(64, 50, 10, 100, seed=3), np.float32, 1e-5, 1e-4),
    ("c2-mini-f64", lambda: P.random_feasible(64, 20, 5, 40, seed=2), np.float64, 1e-3, 1e-9),
    ("c2-mini-f32", lambda: P.random_feasible(64, 20, 5, 40, seed=2), np.float32, 1e-3, 1e-4),
    ("c1-nnls-f64", lambda: P.nnls(256, seed=1), np.float64, 1e-10, 1e-9),
    ("c1-nnls-f32", lambda: P.nnls(256, seed=1), np.float32, 1e-5, 1e-4),
    # sizes served by the workgroup-per-QP matrix-core kernel.  At n >= 70 the reduced KKT
    # matrices reach condition numbers near 1e7 at the stopping point, so last-bit differences of
    # an fp64 elimination move x by up to ~1e-16 * 1e7: the fp64 bar there is 3e-9.
    ("mid-f64", lambda: P.random_feasible(32, 100, 20, 250, seed=6), np.float64, 1e-5, 3e-9),
    ("mid-f32", lambda: P.random_feasible(32, 100, 20, 250, seed=6), np.float32, 1e-5, 1e-4),
    ("big-f64", lambda: P.random_feasible(16, 128, 32, 384, seed=7), np.float64, 1e-5, 3e-9),
    ("odd-f64", lambda: P.random_feasible(16, 70, 3, 140, seed=8), np.float64, 1e-5, 3e-9),
    ("tall-f64", lambda: P.random_feasible(16, 20, 2, 300, seed=9), np.float64, 1e-5, 3e-9),
]


def _run(name, gen, dt, tol, mode):
    import qpax_amd

    data = [np.asarray(v, dt) for v in gen()]
    print(f"[case] {name} mode={mode} shapes={[d.shape for d in data]}", flush=True)
    dev = [torch.as_tensor(v, device="cuda") for v in data]
    out = qpax_amd.solve_qp(*dev, solver_tol=tol, mode=mode)
    torch.cuda.synchronize()
    return data, [to_np(t) for t in out]


@pytest.mark.parametrize("name,gen,dt,tol,xtol", CASES, ids=[c[0] for c in CASES])
def test_solve_matches_oracle(name, gen, dt, tol, xtol):
    data, (x, s, z, y, conv, iters) = _run(name, gen, dt, tol, "fused")
    ref = O.solve_qp(*[d.astype(np.float64) for d in data], solver_tol=tol)
    xo, so, zo, yo, co, io = ref
    assert x.dtype == dt
    ok = co == 1
    assert ok.mean() > 0.9
    np.testing.assert_array_equal(conv, co)
    assert np.abs(iters - io).max() <= 1
    ex = rel_err(x[ok], xo[ok])
    print(f"{name}: x rel err max {ex.max():.2e} median {np.median(ex):.2e}")
    assert ex.max() < xtol
    for a, b in ((s, so), (z, zo)):
        assert rel_err(a[ok], b[ok]).max() < max(xtol, 1e-6) * 10


@pytest.mark.parametrize("name,gen,dt,tol,xtol", CASES, ids=[c[0] for c in CASES])
def test_plain_matches_oracle(name, gen, dt, tol, xtol):
    """The plain per-stage path (a different elimination order than the fused one) meets the same
    bar against the oracle."""
    data, (x, s, z, y, conv, iters) = _run(name, gen, dt, tol, "plain")
    xo, so, zo, yo, co, io = O.solve_qp(*[d.astype(np.float64) for d in data], solver_tol=tol)
    ok = co == 1
    np.testing.assert_array_equal(conv, co)
    assert np.abs(iters - io).max() <= 1
    assert rel_err(x[ok], xo[ok]).max() < xtol


@pytest.mark.parametrize("dt,B", [(torch.float32, 2304), (torch.float64, 256)], ids=["f32-2304", "f64-256"])
def test_relaxed_graph_replays_match_eager(dt, B):
    """A uniform solve + relax runs the relaxations of its last problems (the last 8 per CU) as
    separate work items (ready flags and a ticket counter in the workspace, reset by a kernel at
    each launch), the others in their solve's item: captured in a hipGraph and replayed, with
    unrelated work in between, every replay gives the eager result bit for bit, and that result
    is the oracle's.  B = 2304 mixes both kinds of item on a 256-CU part; B = 256 splits all."""
    from qpax_amd._tensors import normalize
    from qpax_amd.pdip import solve_qp_relaxed_batch

    data = P.random_feasible(B, 50, 10, 100, seed=21)
    print(f"[case] relaxed graph replay dt={dt} B={B} n=50", flush=True)
    bt = normalize(*[torch.tensor(np.asarray(v), dtype=dt, device="cuda") for v in data])
    eager = [t.clone() for t in solve_qp_relaxed_batch(bt, 1e-5, 1e-3)]
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        solve_qp_relaxed_batch(bt, 1e-5, 1e-3)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        out = solve_qp_relaxed_batch(bt, 1e-5, 1e-3)
    for rep in range(3):
        for t in out:
            t.zero_()
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip(out, eager):
            assert torch.equal(a, b), rep
        (torch.randn(1024, device="cuda") * 2).sum().item()
    npdt = np.float32 if dt == torch.float32 else np.float64
    d64 = [np.asarray(np.asarray(v, npdt), np.float64) for v in data]
    xo, rel, info = O.solve_qp_primal_fwd(*d64, solver_tol=1e-5, target_kappa=1e-3)
    ok = (info[0] == 1) & (info[2] == 1)
    assert ok.mean() > 0.9
    xr = to_np(eager[6]).reshape(B, 50).astype(np.float64)
    assert rel_err(xr[ok], rel[0][ok]).max() < (1e-7 if dt == torch.float64 else 1e-4)
