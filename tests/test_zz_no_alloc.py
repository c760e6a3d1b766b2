"""qpx_solve_qp_relaxed_noalloc (the entry point that allocates nothing: its scratch is a device
array inside the library) against qpx_solve_qp_relaxed in fused mode with a caller workspace: the
same kernels, so every output bit for bit, on two calls in a row, the second with a smaller batch
(the scratch keeps the first call's contents), and a clear error beyond the scratch size."""

import ctypes

import numpy as np
import pytest

from tests import problems as P

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _outs(B, n, me, mi, dt):
    f = lambda k: torch.empty(B * k, dtype=dt, device="cuda")  # noqa: E731
    i = lambda: torch.empty(B, dtype=torch.int32, device="cuda")  # noqa: E731
    return [f(n), f(mi), f(mi), f(me), i(), i(), f(n), f(mi), f(mi), f(me), i(), i()]


def _p(t):
    return t.data_ptr() if t.numel() else None


def _call(lib, noalloc, data, dt, tol, kappa):
    from qpax_amd._tensors import dtype_code, stream_of

    Q, q, A, b, G, h = data
    B, n, me, mi = Q.shape[0], Q.shape[1], A.shape[1], G.shape[1]
    out = _outs(B, n, me, mi, dt)
    ops = [_p(t) for t in data]
    st_ptr = stream_of(Q.device)
    if noalloc:
        st = lib.qpx_solve_qp_relaxed_noalloc(dtype_code(dt), B, n, me, mi, *ops, tol, kappa, *[_p(t) for t in out],
                                              st_ptr)
    else:
        nb = lib.qpx_workspace_size(B, n, me, mi)
        ws = torch.empty(nb, dtype=torch.uint8, device="cuda")
        st = lib.qpx_solve_qp_relaxed(dtype_code(dt), 1, B, n, me, mi, *ops, tol, kappa, *[_p(t) for t in out],
                                      _p(ws), nb, st_ptr)
    torch.cuda.synchronize()
    return st, [t.cpu().numpy() for t in out]


@pytest.mark.parametrize("dt", [torch.float32, torch.float64], ids=["f32", "f64"])
def test_noalloc_matches_workspace_entry_point(dt):
    from qpax_amd import _lib

    lib = _lib.load()
    # first call: the headline shape (wavefront kernel, split relaxations); second: a smaller batch
    # of another size (the scratch still holds the first call's flags and state)
    for B, n, me, mi, seed in ((4096, 50, 10, 100, 3), (300, 20, 5, 40, 4)):
        print(f"[case] noalloc vs workspace dt={dt} B={B} n={n}", flush=True)
        data = [torch.tensor(np.asarray(v), dtype=dt, device="cuda") for v in P.random_feasible(B, n, me, mi, seed=seed)]
        st1, ref = _call(lib, False, data, dt, 1e-5, 1e-3)
        st2, got = _call(lib, True, data, dt, 1e-5, 1e-3)
        assert st1 == 0 and st2 == 0, lib.qpx_last_error()
        for k, (a, b) in enumerate(zip(got, ref)):
            np.testing.assert_array_equal(a, b, err_msg=f"output {k} B={B}")
        assert (ref[4] == 1).mean() > 0.9


def test_noalloc_rejects_sizes_beyond_its_scratch():
    from qpax_amd import _lib
    from qpax_amd._tensors import stream_of

    lib = _lib.load()
    cap = lib.qpx_noalloc_scratch_bytes()
    assert cap == 256 << 20
    dummy = torch.empty(1, device="cuda")
    p = dummy.data_ptr()
    # 20000 problems of n=50, m_eq=10, m_ineq=100 need 20000 x 36 KB > 256 MiB: refused before any launch
    st = lib.qpx_solve_qp_relaxed_noalloc(0, 20000, 50, 10, 100, *([p] * 6), 1e-5, 1e-3, *([p] * 12),
                                          stream_of(dummy.device))
    assert st == 4
    assert b"beyond" in lib.qpx_last_error()
