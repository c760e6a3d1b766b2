"""The library's kernels run on the CPU emulator (tests/cpu_emu): bounds and parity without a GPU.

libqpx's HIP sources are compiled for x86-64 against tests/cpu_emu/hip/hip_runtime.h with
AddressSanitizer.  Every thread of a workgroup is a fiber, the wavefront-wide operations are
collectives of the fibers of one wavefront, every device buffer is an exact-size allocation and
each workgroup's dynamic LDS a fresh allocation of exactly the launch's byte count, filled with
NaN bytes (the hardware does not clear LDS either).  A kernel that reads or writes outside any
buffer or outside its LDS image ends the child process with an AddressSanitizer report; a kernel
whose results depend on LDS it did not write produces NaN.  Results are compared with the float64
oracle at the bars of the GPU parity tests.
"""

from __future__ import annotations

import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import qp_oracle as O
from tests import problems as P
from tests._util import rel_err

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "cpu_emu")
sys.path.insert(0, EMU)


@pytest.fixture(scope="module")
def emu():
    if not os.path.exists("/opt/rocm/llvm/bin/clang++"):
        pytest.skip("needs /opt/rocm/llvm/bin/clang++")
    import build_emu

    try:
        rt = build_emu.asan_runtime()
    except RuntimeError as exc:
        pytest.skip(str(exc))
    build_emu.build()
    return rt


def _run(rt, case, tmp_path, tag, **env):
    cin, cout = str(tmp_path / f"{tag}_in.npz"), str(tmp_path / f"{tag}_out.npz")
    np.savez(cin, **case)
    full = dict(os.environ, LD_PRELOAD=rt, ASAN_OPTIONS="detect_leaks=0", **{k: str(v) for k, v in env.items()})
    res = subprocess.run([sys.executable, os.path.join(EMU, "run_emu.py"), cin, cout], env=full,
                         capture_output=True, text=True, timeout=600)
    assert res.returncode == 0, f"emulated run failed (exit {res.returncode}):\n{res.stdout[-2000:]}\n{res.stderr[-6000:]}"
    return dict(np.load(cout))


def _check(tag, data, out, dt, tol, kappa, B, shapes):
    """Forward (x, s, z, y), relaxed point and the six cotangents against the oracle."""
    xo, rel, info = O.solve_qp_primal_fwd(*data, solver_tol=tol, target_kappa=kappa)
    so_, zo_, yo_ = O.solve_qp(*data, solver_tol=tol)[1:4]
    go = O.solve_qp_primal_bwd(*data, rel, 2 * (xo - 1))
    conv, iters = out["conv"], out["iters"]
    np.testing.assert_array_equal(conv, info[0])
    assert np.abs(iters - info[1]).max() <= 1
    ok = (info[0] == 1) & (info[2] == 1)
    assert ok.any()
    same = ok & (iters == info[1]) & (out["riters"] == info[3])
    xtol = 1e-9 if dt == np.float64 else 1e-4
    gtol = 1e-7 if dt == np.float64 else 1e-3
    n, me, mi = shapes
    errs = {}
    for name, ref, size in (("x", xo, n), ("s", so_, mi), ("z", zo_, mi), ("y", yo_, me)):
        if size:
            errs[name] = rel_err(out[name].reshape(B, size)[ok], ref[ok]).max()
    for k, name in enumerate(("dQ", "dq", "dA", "db", "dG", "dh")):
        if go[k].size and same.any():
            errs[name] = rel_err(out[name].reshape(go[k].shape)[same], go[k][same]).max()
    print(f"{tag}: " + " ".join(f"{k} {v:.1e}" for k, v in errs.items()))
    for k, v in errs.items():
        assert v < (gtol if k.startswith("d") else xtol), (tag, k, v)


# one uniform case per wavefront-kernel instance (NMAX, MIC), with odd n, m_eq = 0, misaligned G
# rows (odd m_ineq * n) and m >= 56 (the branch-free backward sweep) among them
WAVE_CASES = [
    ("w32x1-f64", 20, 5, 40, np.float64),
    ("w32x2-f32-oddn", 21, 4, 100, np.float32),
    ("w64x1-f64-m63", 57, 3, 60, np.float64),
    ("w64x2-f32-c3", 50, 10, 100, np.float32),
    ("w64x4-f32-me0-oddn", 43, 0, 201, np.float32),
]


@pytest.mark.parametrize("tag,n,me,mi,dt", WAVE_CASES, ids=[c[0] for c in WAVE_CASES])
@pytest.mark.parametrize("cus", [0, 2], ids=["relax-in-item", "split"])
def test_emulated_wave_uniform(emu, tmp_path, tag, n, me, mi, dt, cus):
    """Uniform fused solve + relax + backward.  With 0 compute units every relaxation runs in its
    solve's work item; with 2, every problem of this 3-problem batch takes the split launch."""
    B, tol, kappa = 3, 1e-6, 1e-3
    data = [np.asarray(v, dt).astype(np.float64) for v in P.random_feasible(B, n, me, mi, seed=n + mi)]
    case = dict(kind="uniform", dtype=np.dtype(dt).name, mode="fused", tol=tol, kappa=kappa,
                **dict(zip("Q q A b G h".split(), data)))
    out = _run(emu, case, tmp_path, tag, EMU_CUS=cus)
    _check(tag, data, out, dt, tol, kappa, B, (n, me, mi))


@pytest.mark.parametrize("n,me,mi,dt", [(66, 6, 90, np.float64), (50, 10, 100, np.float32), (20, 5, 40, np.float64)],
                         ids=["n66-global", "n50-lds-f32", "n20-lds-wave"])
def test_emulated_plain_mode(emu, tmp_path, n, me, mi, dt):
    """The per-stage kernels of the plain mode (host-checked loop): at n >= 64 on the workspace with
    matrix-core G^T D G, below that with the caller's G and the factor staged in LDS (n = 20: one
    wavefront per problem)."""
    B, tol, kappa = 2, 1e-6, 1e-3
    data = [np.asarray(v, dt).astype(np.float64) for v in P.random_feasible(B, n, me, mi, seed=12)]
    case = dict(kind="uniform", dtype=np.dtype(dt).name, mode="plain", tol=tol, kappa=kappa,
                **dict(zip("Q q A b G h".split(), data)))
    out = _run(emu, case, tmp_path, f"plain{n}")
    _check(f"plain-n{n}", data, out, dt, tol, kappa, B, (n, me, mi))


def _ragged_case(probs, dt, tol, kappa, tune=0):
    case = dict(kind="ragged", dtype=dt, tol=tol, kappa=kappa, tune=tune,
                n=[p[1].shape[0] for p in probs], me=[p[3].shape[0] for p in probs],
                mi=[p[5].shape[0] for p in probs])
    for j, k in enumerate("Q q A b G h".split()):
        case[k] = np.concatenate([np.asarray(p[j], np.float64).ravel() for p in probs])
    return case


def _check_ragged(tag, probs, out, dt, tol, kappa, sample):
    off = out["offsets"]
    for i in sample:
        pr = probs[i]
        n, me, mi = pr[1].shape[0], pr[3].shape[0], pr[5].shape[0]
        sl = {"x": (off[i, 1], n), "s": (off[i, 5], mi), "z": (off[i, 5], mi), "y": (off[i, 3], me),
              "xr": (off[i, 1], n), "sr": (off[i, 5], mi), "zr": (off[i, 5], mi), "yr": (off[i, 3], me),
              "dQ": (off[i, 0], n * n), "dq": (off[i, 1], n), "dA": (off[i, 2], me * n), "db": (off[i, 3], me),
              "dG": (off[i, 4], mi * n), "dh": (off[i, 5], mi)}
        one = {k: out[k][o:o + c] for k, (o, c) in sl.items()}
        for k in ("conv", "iters", "rconv", "riters"):
            one[k] = out[k][i:i + 1]
        data = [np.asarray(np.asarray(v, dt), np.float64)[None] for v in pr]
        _check(f"{tag}[{i}] n={n} me={me} mi={mi}", data, one, dt, tol, kappa, 1, (n, me, mi))


def test_emulated_ragged_f32_mixed_kernels(emu, tmp_path):
    """The ragged fp32 batch of the GPU comparison test (40 problems, wavefront, matrix-core and
    generic workgroup kernels, plus two problems beyond the matrix-core kernels): no access
    outside any buffer or LDS image, and every problem against the oracle."""
    probs = P.uneven(40, seed=41, n_lo=8, n_hi=128)
    for sd in (1, 2):
        Q, q, A, b, G, h = P.random_feasible(1, 150, 20, 200, seed=sd)
        probs[3 * sd] = (Q[0], q[0], A[0], b[0], G[0], h[0])
    out = _run(emu, _ragged_case(probs, "float32", 1e-5, 1e-3), tmp_path, "ragged40")
    kinds = {int(b[0]) for b in out["buckets"]}
    assert kinds == {0, 1, 2}
    _check_ragged("ragged40", probs, out, np.float32, 1e-5, 1e-3, range(0, 40, 3))


def test_emulated_ragged_every_launch_shape(emu, tmp_path):
    """qpx_plan_tune runs every candidate shape of every bucket (wavefronts per workgroup 1, 2, 4
    sharing one workgroup's LDS; group kernels of 4 and 8 wavefronts; generic blocks of 64, 128,
    256 threads; persistent queues and one workgroup per problem) on a small fp64 mix."""
    rng = np.random.default_rng(5)
    sizes = [(6, 0, 7), (9, 2, 9), (17, 1, 31), (30, 0, 32), (33, 5, 70), (40, 8, 150), (13, 3, 129),
             (70, 10, 100)]
    probs = []
    for k, (n, me, mi) in enumerate(sizes):
        Q, q, A, b, G, h = P.random_feasible(1, n, me, mi, seed=int(rng.integers(1 << 30)))
        probs.append((Q[0], q[0], A[0], b[0], G[0], h[0]))
    out = _run(emu, _ragged_case(probs, "float64", 1e-6, 1e-3, tune=1), tmp_path, "tune")
    _check_ragged("tune", probs, out, np.float64, 1e-6, 1e-3, range(len(probs)))


def _golden():
    d = np.load(os.path.join(HERE, "golden", "problem_data_pos.npz"))
    iv = np.load(os.path.join(HERE, "golden", "initial_values_pos.npz"))
    return d, iv


def test_emulated_initialize_matches_reference_fixtures(emu, tmp_path):
    """qpx_initialize on the expanded elastic problem and qpx_initialize_elastic on the elastic one
    reproduce the reference's initial_values_pos.npz at its own 1e-6 (test/test_elastic_qp.py:79-95)."""
    d, iv = _golden()
    Q, q, G, h, pen = d["Q"][None], d["q"][None], d["G"][None], d["h"][None], float(d["penalty"])
    Q2, q2, A2, b2, G2, h2 = O.expanded_elastic_problem(Q, q, G, h, pen)
    out = _run(emu, dict(kind="init", dtype="float64", Q=Q2, q=q2, A=A2, b=b2, G=G2, h=h2), tmp_path, "init")
    assert np.linalg.norm(out["x"] - np.concatenate([iv["x"], iv["t"]])) < 1e-6
    assert np.linalg.norm(out["s"] - np.concatenate([iv["s1"], iv["s2"]])) < 1e-6
    assert np.linalg.norm(out["z"] - np.concatenate([iv["z1"], iv["z2"]])) < 1e-6
    A0, b0 = np.zeros((1, 0, 5)), np.zeros((1, 0))
    oe = _run(emu, dict(kind="init_elastic", dtype="float64", Q=Q, q=q, A=A0, b=b0, G=G, h=h, penalty=np.array([pen])),
              tmp_path, "init_el")
    for k in ("x", "t", "s1", "s2", "z1", "z2"):
        assert np.linalg.norm(oe[k] - iv[k]) < 1e-6, k


@pytest.mark.parametrize("debug_step", [0, 1])
def test_emulated_trace_matches_oracle(emu, tmp_path, debug_step):
    """qpx_solve_qp_trace: per-iteration residual norms and step lengths against the oracle's
    solve_qp_trace, for solve_qp's step rule and the reference debug loop's."""
    data = list(P.random_feasible(4, 20, 5, 40, seed=2))
    out = _run(emu, dict(kind="trace", dtype="float64", tol=1e-6, debug_step=debug_step,
                         **dict(zip("Q q A b G h".split(), data))), tmp_path, f"trace{debug_step}")
    (xo, so, zo, yo, co, io), tro = O.solve_qp_trace(*data, solver_tol=1e-6, debug_step=bool(debug_step))
    np.testing.assert_array_equal(out["iters"], io)
    np.testing.assert_array_equal(out["conv"], co)
    tr = out["trace"]
    assert np.array_equal(np.isnan(tr), np.isnan(tro))
    # residual norms to 1e-6 relative above a floor of 1e-9 times the problem's first residual
    # (the last iterations reach rounding-level values, where both sides are noise)
    floor = 1e-9 * np.nanmax(tro[:, 0, :4], axis=1)[:, None, None]
    err = np.abs(tr - tro) - (1e-6 * np.abs(tro) + floor)
    assert np.nanmax(err[..., :4]) <= 0.0
    assert np.nanmax(np.abs(tr[..., 4] - tro[..., 4])) < 1e-9


# box rows (+-e_j, the G = -I of NNLS and the [G_dense; I; -I] of QPTEST): the dense streams stop
# at the last general row and the box rows enter as a diagonal, a gather and single entries
BOX_CASES = [
    ("nnls-f64", lambda: P.nnls(3, seed=1), np.float64, 1e-8),
    ("box20-me3-f32", lambda: P.box_qp(3, 20, 3, 10, seed=5), np.float32, 1e-6),
    ("box40-dense0-f64", lambda: P.box_qp(3, 40, 4, 0, seed=6), np.float64, 1e-6),
    ("box50-me8-f64", lambda: P.box_qp(3, 50, 8, 20, seed=7), np.float64, 1e-6),
]


@pytest.mark.parametrize("tag,gen,dt,tol", BOX_CASES, ids=[c[0] for c in BOX_CASES])
def test_emulated_box_rows(emu, tmp_path, tag, gen, dt, tol):
    data = [np.asarray(v, dt).astype(np.float64) for v in gen()]
    B, n, me, mi = data[0].shape[0], data[0].shape[1], data[2].shape[1], data[4].shape[1]
    case = dict(kind="uniform", dtype=np.dtype(dt).name, mode="fused", tol=tol, kappa=1e-3,
                **dict(zip("Q q A b G h".split(), data)))
    out = _run(emu, case, tmp_path, tag, EMU_CUS=1)
    _check(tag, data, out, dt, tol, 1e-3, B, (n, me, mi))


def test_emulated_box_rows_with_infinite_bounds(emu, tmp_path):
    """Upper bounds of +inf: those rows are zeroed like every row with infinite h, and as zero rows
    they sit among the box rows; the forward matches the oracle (the relaxation of such rows is
    NaN in qpax, so only the solve is compared)."""
    data = list(P.box_qp(3, 24, 2, 6, seed=8, inf_frac=0.3))
    assert np.isinf(data[5]).any()
    B, n, me, mi = 3, 24, 2, data[4].shape[1]
    case = dict(kind="uniform", dtype="float64", mode="fused", tol=1e-6, kappa=1e-3,
                **dict(zip("Q q A b G h".split(), data)))
    out = _run(emu, case, tmp_path, "boxinf", EMU_CUS=1)
    xo, so, zo, yo, co, io = O.solve_qp(*data, solver_tol=1e-6)
    np.testing.assert_array_equal(out["conv"], co)
    assert np.abs(out["iters"] - io).max() <= 1
    assert rel_err(out["x"].reshape(B, n), xo).max() < 1e-9
    fin = np.isfinite(so)
    np.testing.assert_array_equal(np.isinf(out["s"].reshape(B, mi)), ~fin)
    assert np.abs(out["z"].reshape(B, mi)[~fin]).max() == 0.0


@pytest.mark.parametrize("n,me,mi", [(70, 3, 90), (100, 0, 130), (100, 5, 130)],
                         ids=["n70-128thr", "n100-256thr", "n100-256thr-eq"])
def test_emulated_elastic_wide(emu, tmp_path, n, me, mi):
    """The elastic kernels at matrix-core sizes (column sums through the G^T D G chunk's LDS, tiled
    factorisation, row products) in both modes: every access in bounds, forward and gradients
    against the oracle (the consistent relaxation: the reference's flipped signs amplify rounding
    differences about 2x per relaxation step)."""
    if me:
        Q, q, A, b, G, h, _ = P.elastic_infeasible(1, n, mi, frac=1.0, seed=11, me=me)
    else:
        Q, q, G, h, _ = P.elastic_infeasible(1, n, mi, frac=1.0, seed=11)
        A, b = np.zeros((1, 0, n)), np.zeros((1, 0))
    pen = np.array([14.0])
    for mode in ("fused", "plain"):
        out = _run(emu, dict(kind="elastic", dtype="float64", mode=mode, Q=Q, q=q, A=A, b=b, G=G, h=h, penalty=pen,
                             tol=1e-6, kappa=1e-3, flags=2), tmp_path, f"el_{mode}")
        kw = dict(A=A, b=b) if me else {}
        xo, rel, info = O.solve_qp_elastic_primal_fwd(Q, q, G, h, pen, solver_tol=1e-6, target_kappa=1e-3,
                                                      consistent=True, **kw)
        assert out["conv"][0] == info[0][0] == 1
        assert abs(int(out["iters"][0]) - int(info[1][0])) <= 1
        assert rel_err(out["x"], xo.ravel()).max() < 1e-8
        go = O.solve_qp_elastic_primal_bwd(Q, q, G, h, rel, 2 * (xo - 1), **({"A": A} if me else {}))
        for name, g in zip(("dQ", "dq", "dG", "dh"), go[:4]):
            e = rel_err(out[name], np.ravel(g)).max()
            print(f"{mode} {name}: {e:.2e}")
            assert e < 1e-6, name
