"""qpax.solve_qp on the HIP library -- host side of qpax/pdip.py.

Same arguments and return tuple as the reference: ``solve_qp(Q, q, A, b, G, h, solver_tol=1e-3)
-> (x, s, z, y, converged, iters)`` (pdip.py:181-251), the semi-public ``initialize(Q, q, A, b, G,
h) -> (x, s, z, y)`` (pdip.py:15-35) and ``solve_qp_debug`` (pdip.py:317-360, per-iteration table
from a device-side trace; ``solve_qp_trace`` returns the trace itself).  Inputs are torch tensors
(or arrays) with any leading batch axes; the whole batch is solved by one library call.  Arithmetic
is float64 inside the kernels for either input dtype; outputs have the input dtype, flags are int32.
"""

from __future__ import annotations

import torch

from . import _lib
from ._tensors import Batch, dtype_code, normalize, ptr, stream_of


def _workspace(bt: Batch):
    nbytes = _lib.load().qpx_workspace_size(bt.B, bt.n, bt.me, bt.mi)
    return torch.empty(max(nbytes, 1), dtype=torch.uint8, device=bt.device), nbytes


def _special(bt: Batch):
    """Static shape branches of pdip.py:208-217 (no inequality rows)."""
    from .special import solve_eq_only, solve_unconstrained

    if bt.me == 0:
        x = solve_unconstrained(bt)
        y = bt.out(0)
    else:
        x, y = solve_eq_only(bt)
    e = bt.out(0)
    B = bt.B
    one = torch.ones(B, dtype=torch.int32, device=bt.device)
    zero = torch.zeros(B, dtype=torch.int32, device=bt.device)
    return (bt.unflatten(x, bt.n), bt.unflatten(e, 0), bt.unflatten(e, 0), bt.unflatten(y, bt.me),
            bt.unflatten(one), bt.unflatten(zero))


def solve_qp_batch(bt: Batch, solver_tol=1e-3, mode="fused"):
    """Forward solve of a normalised batch; returns flat (B, ...) tensors."""
    lib = _lib.load()
    x, s, z, y = bt.out(bt.n), bt.out(bt.mi), bt.out(bt.mi), bt.out(bt.me)
    conv = torch.empty(bt.B, dtype=torch.int32, device=bt.device)
    iters = torch.empty(bt.B, dtype=torch.int32, device=bt.device)
    ws, nbytes = _workspace(bt)
    st = lib.qpx_solve_qp(dtype_code(bt.dtype), _lib.MODES[mode], bt.B, bt.n, bt.me, bt.mi,
                          *[ptr(bt[k]) for k in bt.names], float(solver_tol),
                          ptr(x), ptr(s), ptr(z), ptr(y), ptr(conv), ptr(iters),
                          ptr(ws), nbytes, stream_of(bt.device))
    _lib.check(st, "qpx_solve_qp")
    return x, s, z, y, conv, iters


def solve_qp_relaxed_batch(bt: Batch, solver_tol=1e-5, target_kappa=1e-3, mode="fused", check=True):
    """Solve then relax (diff_qp.py:61-71); returns (x, s, z, y, conv, iters, xr, sr, zr, yr,
    rconv, riters) as flat tensors.  With `check`, raises if a relaxation did not start
    (riters = -1, see check_relaxed); that reads the flags back, i.e. synchronises."""
    lib = _lib.load()
    x, s, z, y = bt.out(bt.n), bt.out(bt.mi), bt.out(bt.mi), bt.out(bt.me)
    xr, sr, zr, yr = bt.out(bt.n), bt.out(bt.mi), bt.out(bt.mi), bt.out(bt.me)
    flags = torch.empty(4, bt.B, dtype=torch.int32, device=bt.device)
    ws, nbytes = _workspace(bt)
    st = lib.qpx_solve_qp_relaxed(dtype_code(bt.dtype), _lib.MODES[mode], bt.B, bt.n, bt.me, bt.mi,
                                  *[ptr(bt[k]) for k in bt.names], float(solver_tol), float(target_kappa),
                                  ptr(x), ptr(s), ptr(z), ptr(y), ptr(flags[0]), ptr(flags[1]),
                                  ptr(xr), ptr(sr), ptr(zr), ptr(yr), ptr(flags[2]), ptr(flags[3]),
                                  ptr(ws), nbytes, stream_of(bt.device))
    _lib.check(st, "qpx_solve_qp_relaxed")
    if check:
        check_relaxed(flags[3])
    return x, s, z, y, flags[0], flags[1], xr, sr, zr, yr, flags[2], flags[3]


def check_relaxed(riters):
    """Raises if a relaxation reported riters = -1: its wait for its own solve gave up (the library
    then writes NaN to that problem's relaxed point).  Reads the flags back, so it is skipped while
    a graph is being captured."""
    if riters.numel() == 0 or torch.cuda.is_current_stream_capturing():
        return
    bad = int((riters < 0).sum())
    if bad:
        raise RuntimeError(f"qpx_solve_qp_relaxed: the relaxation of {bad} problem(s) did not start (riters = -1); "
                           "their relaxed points are NaN")


def solve_qp(Q, q, A, b, G, h, solver_tol: float = 1e-3, mode: str = "fused"):
    """Solve min 1/2 x'Qx + q'x s.t. Ax = b, Gx <= h (qpax/pdip.py:181).

    Returns (x, s, z, y, converged, iters).  ``mode`` selects the fused single-launch path or the
    plain per-stage path; they run different eliminations and agree to rounding."""
    bt = normalize(Q, q, A, b, G, h)
    if bt.mi == 0:
        return _special(bt)
    x, s, z, y, conv, iters = solve_qp_batch(bt, solver_tol, mode)
    return (bt.unflatten(x, bt.n), bt.unflatten(s, bt.mi), bt.unflatten(z, bt.mi), bt.unflatten(y, bt.me),
            bt.unflatten(conv), bt.unflatten(iters))


def initialize(Q, q, A, b, G, h):
    """qpax.pdip.initialize (pdip.py:15-35) -> (x, s, z, y): the CVXOPT initial point of the
    predictor-corrector loop, on the problem as solve_qp hands it over (Q symmetrised, rows with
    infinite h zeroed, pdip.py:220-223)."""
    bt = normalize(Q, q, A, b, G, h)
    if bt.mi == 0:
        raise ValueError("initialize needs inequality rows (m_ineq >= 1), as the reference's does")
    lib = _lib.load()
    x, s, z, y = bt.out(bt.n), bt.out(bt.mi), bt.out(bt.mi), bt.out(bt.me)
    ws, nbytes = _workspace(bt)
    st = lib.qpx_initialize(dtype_code(bt.dtype), bt.B, bt.n, bt.me, bt.mi, *[ptr(bt[k]) for k in bt.names],
                            ptr(x), ptr(s), ptr(z), ptr(y), ptr(ws), nbytes, stream_of(bt.device))
    _lib.check(st, "qpx_initialize")
    return bt.unflatten(x, bt.n), bt.unflatten(s, bt.mi), bt.unflatten(z, bt.mi), bt.unflatten(y, bt.me)


def solve_qp_trace(Q, q, A, b, G, h, solver_tol: float = 1e-3, debug_step: bool = False):
    """solve_qp with a per-iteration trace: returns ((x, s, z, y, converged, iters), trace) where
    trace[..., k, :] holds the infinity norms of r1, r2, r3, r4 at iteration k's entry and the step
    length it took (NaN rows after the last iteration).  ``debug_step`` uses the step length rule
    of the reference's debug loop (pdip.py:281) instead of solve_qp's (pdip.py:112)."""
    bt = normalize(Q, q, A, b, G, h)
    if bt.mi == 0:
        raise ValueError("the trace covers the interior-point loop (m_ineq >= 1)")
    lib = _lib.load()
    x, s, z, y = bt.out(bt.n), bt.out(bt.mi), bt.out(bt.mi), bt.out(bt.me)
    conv = torch.empty(bt.B, dtype=torch.int32, device=bt.device)
    iters = torch.empty(bt.B, dtype=torch.int32, device=bt.device)
    trace = torch.empty(bt.B, _lib.TRACE_ITERS, _lib.TRACE_COLS, dtype=torch.float64, device=bt.device)
    ws, nbytes = _workspace(bt)
    st = lib.qpx_solve_qp_trace(dtype_code(bt.dtype), bt.B, bt.n, bt.me, bt.mi, *[ptr(bt[k]) for k in bt.names],
                                float(solver_tol), _lib.QPX_TRACE_DEBUG_STEP if debug_step else 0, ptr(x), ptr(s),
                                ptr(z), ptr(y), ptr(conv), ptr(iters), ptr(trace), ptr(ws), nbytes,
                                stream_of(bt.device))
    _lib.check(st, "qpx_solve_qp_trace")
    out = (bt.unflatten(x, bt.n), bt.unflatten(s, bt.mi), bt.unflatten(z, bt.mi), bt.unflatten(y, bt.me),
           bt.unflatten(conv), bt.unflatten(iters))
    return out, trace.reshape(*bt.batch_shape, _lib.TRACE_ITERS, _lib.TRACE_COLS)


def solve_qp_debug(Q, q, A, b, G, h, solver_tol: float = 1e-3):
    """qpax.pdip.solve_qp_debug (pdip.py:317-360): solve with the debug loop's step rule and print
    the per-iteration table (iteration, |r1|, |r2|, |r3|, |r4|, alpha) of each problem."""
    bt = normalize(Q, q, A, b, G, h)
    if bt.mi == 0:
        return solve_qp(Q, q, A, b, G, h, solver_tol)
    out, trace = solve_qp_trace(Q, q, A, b, G, h, solver_tol, debug_step=True)
    tr = trace.reshape(bt.B, _lib.TRACE_ITERS, _lib.TRACE_COLS).cpu().numpy()
    its = out[5].reshape(bt.B).cpu().numpy()
    for p in range(bt.B):
        if bt.B > 1:
            print(f"problem {p}")
        print("iter      r1          r2         r3         r4        alpha")
        print("------------------------------------------------------")
        for k in range(int(its[p])):
            r1, r2, r3, r4, al = tr[p, k]
            print("%3d   %9.2e   %9.2e  %9.2e  %9.2e   % 6.4f" % (k, r1, r2, r3, r4, al))
    return out
