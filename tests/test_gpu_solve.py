"""GPU parity of the forward solve and the relaxation against the float64 oracle."""

import numpy as np
import pytest

from oracle import qp_oracle as O
from tests import problems as P
from tests._util import conditioning_floor, rel_err, to_np

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
    # sizes served by the workgroup-per-QP matrix-core kernel
    ("mid-f64", lambda: P.random_feasible(32, 100, 20, 250, seed=6), np.float64, 1e-5, 1e-9),
    ("mid-f32", lambda: P.random_feasible(32, 100, 20, 250, seed=6), np.float32, 1e-5, 1e-4),
    ("big-f64", lambda: P.random_feasible(16, 128, 32, 384, seed=7), np.float64, 1e-5, 1e-9),
    ("odd-f64", lambda: P.random_feasible(16, 70, 3, 140, seed=8), np.float64, 1e-5, 1e-9),
    ("tall-f64", lambda: P.random_feasible(16, 20, 2, 300, seed=9), np.float64, 1e-5, 1e-9),
    # one uniform batch per wavefront-kernel instance (NMAX, MIC) beyond (32, 1) and (64, 2), with
    # odd n, m_eq = 0, n = 17 and n = 49, 51, 52 (G reaching 1..4 columns of the last tile column:
    # the 4x4x4 strip products), m >= 56 (the branch-free backward sweep) and m_ineq * n odd
    # (G rows of odd problems not 16-byte aligned: the scalar G copy)
    ("w32x2-f32", lambda: P.random_feasible(64, 21, 4, 100, seed=31), np.float32, 1e-5, 1e-4),
    ("w32x2-f64", lambda: P.random_feasible(64, 21, 4, 100, seed=31), np.float64, 1e-6, 1e-9),
    ("w32x1-strip-f64", lambda: P.random_feasible(64, 17, 3, 41, seed=32), np.float64, 1e-6, 1e-9),
    ("w64x1-f32", lambda: P.random_feasible(64, 40, 8, 48, seed=33), np.float32, 1e-5, 1e-4),
    ("w64x1-m63-f64", lambda: P.random_feasible(64, 57, 3, 60, seed=34), np.float64, 1e-6, 1e-9),
    ("w64x1-strip49-f64", lambda: P.random_feasible(64, 49, 4, 63, seed=35), np.float64, 1e-6, 1e-9),
    ("w64x2-strip51-me0-f32", lambda: P.random_feasible(64, 51, 0, 120, seed=36), np.float32, 1e-5, 1e-4),
    ("w64x4-f32", lambda: P.random_feasible(64, 40, 8, 200, seed=37), np.float32, 1e-5, 1e-4),
    ("w64x4-odd-me0-f64", lambda: P.random_feasible(64, 43, 0, 201, seed=38), np.float64, 1e-6, 1e-9),
    ("w64x4-strip52-f64", lambda: P.random_feasible(64, 52, 8, 256, seed=39), np.float64, 1e-6, 1e-9),
    # box rows below general rows (G = [G_dense; I; -I], the layout of the reference's QPTEST):
    # the dense streams stop at the last general row, the +-e_j rows enter as a diagonal and gathers
    ("box20-f32", lambda: P.box_qp(64, 20, 3, 10, seed=5), np.float32, 1e-5, 1e-4),
    ("box40-dense0-f64", lambda: P.box_qp(64, 40, 4, 0, seed=6), np.float64, 1e-6, 1e-9),
    ("box50-f64", lambda: P.box_qp(64, 50, 8, 20, seed=7), np.float64, 1e-6, 1e-9),
    ("box24-inf-f64", lambda: P.box_qp(64, 24, 2, 6, seed=8, inf_frac=0.3), np.float64, 1e-6, 1e-9),
]


def _assert_state(name, mode, dt, xtol, ok, got, oref, data, tol):
    """x, s, z and y of the converged problems within the bar (relative 2-norm per problem).

    float32 inputs: 1e-4.  float64 inputs: 1e-9, or four times the conditioning floor of the batch
    where that is larger -- the spread of the float64 oracle itself when every input entry moves
    by one rounding (tests._util.conditioning_floor).  Near the stopping point the reduced KKT
    systems of these problems are ill-conditioned enough that two correct float64 eliminations
    (this library's LDL^T or Cholesky, the oracle's QR) end 1e-9 to 3e-8 apart in s, z and y; the
    floor measures that per batch, so the bar tracks it instead of a constant."""
    errs = {}
    for k, a, b in zip("xszy", got, oref[:4]):
        if b.shape[-1]:
            errs[k] = rel_err(a[ok], b[ok]).max()
    floor = {}
    if dt == np.float64:
        d64 = [d.astype(np.float64) for d in data]
        floor = conditioning_floor(lambda *v: O.solve_qp(*v, solver_tol=tol), d64, oref, ok)
    print(f"{name} {mode}: " + ", ".join(f"{k} rel err max {v:.2e} (floor {floor.get(k, 0.0):.1e})"
                                           for k, v in errs.items()))
    for k, v in errs.items():
        assert v < max(xtol, 4.0 * floor.get(k, 0.0)), (k, v, floor.get(k))


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
    _assert_state(name, "fused", dt, xtol, ok, (x, s, z, y), ref, data, tol)


@pytest.mark.parametrize("name,gen,dt,tol,xtol", CASES, ids=[c[0] for c in CASES])
def test_plain_matches_oracle(name, gen, dt, tol, xtol):
    """The plain per-stage path (a different elimination order than the fused one) meets the same
    bar against the oracle."""
    data, (x, s, z, y, conv, iters) = _run(name, gen, dt, tol, "plain")
    ref = O.solve_qp(*[d.astype(np.float64) for d in data], solver_tol=tol)
    xo, so, zo, yo, co, io = ref
    ok = co == 1
    np.testing.assert_array_equal(conv, co)
    assert np.abs(iters - io).max() <= 1
    _assert_state(name, "plain", dt, xtol, ok, (x, s, z, y), ref, data, tol)


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


def test_split_wait_giving_up_is_reported(monkeypatch):
    """A split relaxation whose bounded wait for its solve gives up writes NaN to its relaxed point
    and reports riters = -1, and the host raises.  The wait is made to give up at once from the
    host (QPX_SPLIT_SPIN=0: zero polls), so every kernel still ends normally."""
    import qpax_amd.pdip as pdip
    from qpax_amd._tensors import normalize

    cus = torch.cuda.get_device_properties(0).multi_processor_count
    B = 8 * cus + 64                      # the last 8 per CU split, the first 64 relax in their item
    F = B - 8 * cus
    data = P.random_feasible(B, 20, 5, 40, seed=41)
    bt = normalize(*[torch.tensor(np.asarray(v), dtype=torch.float64, device="cuda") for v in data])
    print(f"[case] split wait giving up B={B} F={F}", flush=True)
    good = [t.clone() for t in pdip.solve_qp_relaxed_batch(bt, 1e-6, 1e-3)]
    torch.cuda.synchronize()
    monkeypatch.setenv("QPX_SPLIT_SPIN", "0")
    with pytest.raises(RuntimeError, match="riters = -1"):
        pdip.solve_qp_relaxed_batch(bt, 1e-6, 1e-3)
    monkeypatch.setattr(pdip, "check_relaxed", lambda riters: None)
    out = [to_np(t) for t in pdip.solve_qp_relaxed_batch(bt, 1e-6, 1e-3)]
    # under autograd (no read-back in the forward) the problems whose relaxation did not start get
    # NaN cotangents, the others finite ones
    import qpax_amd

    ts = [torch.tensor(np.asarray(v), dtype=torch.float64, device="cuda", requires_grad=True) for v in data]
    x = qpax_amd.solve_qp_primal(*ts, solver_tol=1e-6, target_kappa=1e-3)
    ((x - 1.0) ** 2).sum().backward()
    gq = to_np(ts[1].grad)
    assert np.isnan(gq[F:]).all() and np.isfinite(gq[:F]).all()
    monkeypatch.delenv("QPX_SPLIT_SPIN")
    good = [to_np(t) for t in good]
    xr, sr, zr, yr, rconv, riters = out[6], out[7], out[8], out[9], out[10], out[11]
    assert np.all(riters[F:] == -1) and np.all(rconv[F:] == 0)
    for v in (xr, sr, zr, yr):
        assert np.isnan(v[F:]).all() and np.isfinite(v[:F]).all()
    # the solves and the relaxations that ran in their solve's item are untouched
    for k in range(6):
        np.testing.assert_array_equal(out[k], good[k])
    for k in range(6, 12):
        np.testing.assert_array_equal(out[k][:F], good[k][:F])
    assert np.all(good[11] >= 1)
