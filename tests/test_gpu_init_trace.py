"""GPU parity of the semi-public initialisations and of the per-iteration trace.

* qpax.pdip.initialize (pdip.py:15-35) and qpax.elastic_qp.initialize_elastic (elastic_qp.py:27-47)
  against the reference's initial_values_pos.npz at its own 1e-6 (test/test_elastic_qp.py:79-95),
  and on C2-shaped batches against the oracle's initialize.
* solve_qp_trace / solve_qp_debug (pdip.py:254-360): the per-iteration residual norms and step
  lengths against the oracle's trace on C2 problems (n=20, m_eq=5, m_ineq=40).
"""

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
    iv = np.load(os.path.join(GOLD, "initial_values_pos.npz"))
    return d, iv


def _cuda(*vs, dt=torch.float64):
    return [torch.tensor(np.asarray(v), dtype=dt, device="cuda") for v in vs]


def test_initialize_matches_reference_fixture():
    from qpax_amd.elastic_qp import initialize_elastic
    from qpax_amd.pdip import initialize

    d, iv = _fixture()
    print("[case] initialize / initialize_elastic on the reference fixture", flush=True)
    Q, q, G, h, pen = d["Q"][None], d["q"][None], d["G"][None], d["h"][None], float(d["penalty"])
    Q2, q2, A2, b2, G2, h2 = O.expanded_elastic_problem(Q, q, G, h, pen)
    x, s, z, y = [to_np(t) for t in initialize(*_cuda(Q2[0], q2[0], A2[0], b2[0], G2[0], h2[0]))]
    assert np.linalg.norm(x - np.concatenate([iv["x"], iv["t"]])) < 1e-6
    assert np.linalg.norm(s - np.concatenate([iv["s1"], iv["s2"]])) < 1e-6
    assert np.linalg.norm(z - np.concatenate([iv["z1"], iv["z2"]])) < 1e-6
    out = [to_np(t) for t in initialize_elastic(*_cuda(d["Q"], d["q"], d["G"], d["h"]), pen)]
    for a, k in zip(out, ("x", "t", "s1", "s2", "z1", "z2")):
        assert np.linalg.norm(a - iv[k]) < 1e-6, k


@pytest.mark.parametrize("dt,bar", [(np.float64, 1e-9), (np.float32, 1e-4)], ids=["f64", "f32"])
def test_initialize_batch_matches_oracle(dt, bar):
    from qpax_amd.elastic_qp import initialize_elastic
    from qpax_amd.pdip import initialize

    data = [np.asarray(v, dt) for v in P.random_feasible(64, 20, 5, 40, seed=2)]
    print(f"[case] initialize batch dt={dt.__name__}", flush=True)
    tdt = torch.float64 if dt == np.float64 else torch.float32
    got = [to_np(t) for t in initialize(*_cuda(*data, dt=tdt))]
    d64 = [v.astype(np.float64) for v in data]
    Qs = 0.5 * (d64[0] + np.swapaxes(d64[0], -1, -2))
    ref = O.initialize(Qs, *d64[1:])
    for k, a, b in zip("xszy", got, ref):
        e = rel_err(a, b).max()
        print(f"initialize {k}: rel err max {e:.2e}")
        assert e < bar, k
    pen = 3.0
    gel = [to_np(t) for t in initialize_elastic(*_cuda(data[0], data[1], data[4], data[5], dt=tdt), pen)]
    rel = O.initialize_elastic(d64[0], d64[1], d64[4], d64[5], pen)
    for k, a, b in zip(("x", "t", "s1", "s2", "z1", "z2"), gel, rel):
        e = rel_err(a, b).max()
        print(f"initialize_elastic {k}: rel err max {e:.2e}")
        assert e < bar, k


@pytest.mark.parametrize("debug_step", [False, True], ids=["solve-step", "debug-step"])
def test_trace_matches_oracle(debug_step):
    from qpax_amd.pdip import solve_qp_trace

    data = list(P.random_feasible(64, 20, 5, 40, seed=2))
    print(f"[case] solve_qp_trace debug_step={debug_step}", flush=True)
    out, tr = solve_qp_trace(*_cuda(*data), solver_tol=1e-6, debug_step=debug_step)
    tr = to_np(tr)
    (xo, so, zo, yo, co, io), tro = O.solve_qp_trace(*data, solver_tol=1e-6, debug_step=debug_step)
    np.testing.assert_array_equal(to_np(out[4]), co)
    np.testing.assert_array_equal(to_np(out[5]), io)
    assert np.array_equal(np.isnan(tr), np.isnan(tro))
    floor = 1e-9 * np.nanmax(tro[:, 0, :4], axis=1)[:, None, None]
    err = np.abs(tr - tro) - (1e-6 * np.abs(tro) + floor)
    print(f"trace: worst excess over the bar {np.nanmax(err[..., :4]):.2e}, "
          f"step length max diff {np.nanmax(np.abs(tr[..., 4] - tro[..., 4])):.2e}")
    assert np.nanmax(err[..., :4]) <= 0.0
    assert np.nanmax(np.abs(tr[..., 4] - tro[..., 4])) < 1e-9
    ok = co == 1
    assert rel_err(to_np(out[0])[ok], xo[ok]).max() < 1e-9


def test_solve_qp_debug_prints_the_table(capsys):
    from qpax_amd.pdip import solve_qp_debug

    data = list(P.random_feasible(1, 20, 5, 40, seed=3))
    out = solve_qp_debug(*_cuda(*[v[0] for v in data]), solver_tol=1e-6)
    text = capsys.readouterr().out
    lines = [ln for ln in text.splitlines() if ln[:3].strip().isdigit()]
    assert int(out[5]) == len(lines) and int(out[4]) == 1
    (_, _, _, _, co, io), tro = O.solve_qp_trace(*data, solver_tol=1e-6, debug_step=True)
    assert len(lines) == int(io[0])
    last = [float(v) for v in lines[-1].split()[1:]]
    assert abs(last[4] - tro[0, io[0] - 1, 4]) < 1e-4
