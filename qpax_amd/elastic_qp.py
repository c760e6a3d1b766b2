"""qpax's elastic (soft-constraint) QP on the HIP library -- host side of qpax/elastic_qp.py.

    minimise 1/2 x'Qx + q'x + penalty 1't   s.t.  G x - t <= h,  t >= 0

``solve_qp_elastic(Q, q, G, h, penalty, solver_tol=1e-3) -> (x, t, s1, s2, z1, z2, converged,
iters)`` and the differentiable ``solve_qp_elastic_primal(Q, q, G, h, penalty, solver_tol=1e-5,
target_kappa=1e-3) -> x`` (cotangents for Q, q, G, h; none for the penalty), as the reference
(qpax/elastic_qp.py:184, :420-469).  Optional equality rows ``A, b`` are an extension (the
reference's elastic functions take none); with them ``solve_qp_elastic`` returns y after z2 and the
primal's backward also returns cotangents for A and b.
"""

from __future__ import annotations

import torch

from . import _lib
from ._tensors import dtype_code, normalize, ptr, stream_of
from .diff_qp import _reduce_to
from .pdip import _workspace


def _prep(Q, q, G, h, penalty, A, b):
    qt = q if isinstance(q, torch.Tensor) else torch.as_tensor(q)
    n = qt.shape[-1]
    if A is None:
        A = torch.zeros((0, n), dtype=qt.dtype)
        b = torch.zeros((0,), dtype=qt.dtype)
    bt = normalize(Q, q, A, b, G, h)
    pen = torch.as_tensor(penalty, dtype=bt.dtype, device=bt.device)
    pen = pen.expand(bt.batch_shape).reshape(bt.B).contiguous() if pen.dim() <= len(bt.batch_shape) else pen
    return bt, pen


RELAX_MODES = {"reference": 0, "consistent": _lib.QPX_ELASTIC_CONSISTENT}


def _solve(bt, pen, solver_tol, mode, relax=None, target_kappa=1e-3):
    lib = _lib.load()
    n, me, mi = bt.n, bt.me, bt.mi
    outs = [bt.out(n), bt.out(mi), bt.out(mi), bt.out(mi), bt.out(mi), bt.out(mi), bt.out(me)]
    flags = torch.empty(4, bt.B, dtype=torch.int32, device=bt.device)
    ws, nbytes = _workspace(bt)
    common = (dtype_code(bt.dtype), _lib.MODES[mode], bt.B, n, me, mi, *[ptr(bt[k]) for k in bt.names], ptr(pen))
    if relax is None:
        st = lib.qpx_solve_qp_elastic(*common, float(solver_tol), *[ptr(o) for o in outs], ptr(flags[0]),
                                      ptr(flags[1]), ptr(ws), nbytes, stream_of(bt.device))
        _lib.check(st, "qpx_solve_qp_elastic")
        return outs, flags[0], flags[1], None, None, None
    routs = [bt.out(n), bt.out(mi), bt.out(mi), bt.out(mi), bt.out(mi), bt.out(mi), bt.out(me)]
    st = lib.qpx_solve_qp_elastic_relaxed(*common, float(solver_tol), float(target_kappa), RELAX_MODES[relax],
                                          *[ptr(o) for o in outs], ptr(flags[0]), ptr(flags[1]),
                                          *[ptr(o) for o in routs], ptr(flags[2]), ptr(flags[3]), ptr(ws), nbytes,
                                          stream_of(bt.device))
    _lib.check(st, "qpx_solve_qp_elastic_relaxed")
    return outs, flags[0], flags[1], routs, flags[2], flags[3]


def solve_qp_elastic(Q, q, G, h, penalty, solver_tol: float = 1e-3, A=None, b=None, mode: str = "fused"):
    """qpax/elastic_qp.py:184-228 -> (x, t, s1, s2, z1, z2, converged, iters)."""
    bt, pen = _prep(Q, q, G, h, penalty, A, b)
    outs, conv, iters, *_ = _solve(bt, pen, solver_tol, mode)
    x, t, s1, s2, z1, z2, y = outs
    mi = bt.mi
    res = [bt.unflatten(x, bt.n)] + [bt.unflatten(v, mi) for v in (t, s1, s2, z1, z2)]
    if A is not None:
        res.append(bt.unflatten(y, bt.me))
    return (*res, bt.unflatten(conv), bt.unflatten(iters))


def diff_qp_elastic_batch(bt, pen, relaxed, dl_dx, mode="fused", exact_dG=False):
    lib = _lib.load()
    g = dl_dx.to(bt.dtype).reshape(bt.B, bt.n).contiguous()
    dQ, dq, dA = bt.out(bt.n, bt.n), bt.out(bt.n), bt.out(bt.me, bt.n)
    db, dG, dh = bt.out(bt.me), bt.out(bt.mi, bt.n), bt.out(bt.mi)
    ws, nbytes = _workspace(bt)
    st = lib.qpx_diff_qp_elastic(dtype_code(bt.dtype), _lib.MODES[mode], bt.B, bt.n, bt.me, bt.mi,
                                 *[ptr(bt[k]) for k in bt.names], ptr(pen), *[ptr(r) for r in relaxed], ptr(g),
                                 _lib.QPX_EXACT_DG if exact_dG else 0, ptr(dQ), ptr(dq), ptr(dA), ptr(db), ptr(dG),
                                 ptr(dh), ptr(ws), nbytes, stream_of(bt.device))
    _lib.check(st, "qpx_diff_qp_elastic")
    return dQ, dq, dA, db, dG, dh


class _SolveQPElasticPrimal(torch.autograd.Function):
    @staticmethod
    def forward(ctx, Q, q, G, h, penalty, A, b, solver_tol, target_kappa, mode, relax, exact_dG):
        bt, pen = _prep(Q, q, G, h, penalty, A if A.numel() or A.dim() > 1 else None, b)
        need = any(ctx.needs_input_grad[:4]) or any(ctx.needs_input_grad[5:7])
        if need:
            outs, *_, routs, _, _ = _solve(bt, pen, solver_tol, mode, relax=relax, target_kappa=target_kappa)
            ctx.save_for_backward(*[bt[k] for k in bt.names], pen, *routs)
            ctx.meta = (bt.batch_shape, bt.n, bt.me, bt.mi, mode, exact_dG,
                        [tuple(t.shape) for t in (Q, q, G, h, A, b)])
        else:
            outs, *_ = _solve(bt, pen, solver_tol, mode)
        return bt.unflatten(outs[0], bt.n)

    @staticmethod
    def backward(ctx, gx):
        Q, q, A, b, G, h, pen, *rel = ctx.saved_tensors
        batch_shape, n, me, mi, mode, exact_dG, shapes = ctx.meta
        bt = normalize(Q, q, A, b, G, h)
        dQ, dq, dA, db, dG, dh = diff_qp_elastic_batch(bt, pen, rel, gx.contiguous(), mode, exact_dG)
        need = ctx.needs_input_grad
        out = []
        for gr, tail, shp, nd in zip((dQ, dq, dG, dh, dA, db), [(n, n), (n,), (mi, n), (mi,), (me, n), (me,)], shapes,
                                     (need[0], need[1], need[2], need[3], need[5], need[6])):
            out.append(_reduce_to(gr.reshape(*batch_shape, *tail), shp) if nd else None)
        dQo, dqo, dGo, dho, dAo, dbo = out
        return dQo, dqo, dGo, dho, None, dAo, dbo, None, None, None, None, None


def solve_qp_elastic_primal(Q, q, G, h, penalty, solver_tol: float = 1e-5, target_kappa: float = 1e-3, A=None,
                            b=None, mode: str = "fused", relax: str = "reference", exact_dG: bool = False):
    """Primal solution x of the elastic QP, differentiable w.r.t. (Q, q, G, h) [and (A, b)]
    (qpax/elastic_qp.py:420-469).  ``relax="consistent"`` runs the relaxation with the forward's
    signs of t and r2; the default reproduces the reference's (qpax/elastic_qp.py:257-261)."""
    qt = q if isinstance(q, torch.Tensor) else torch.as_tensor(q)
    if A is None:
        A = torch.zeros((0,), dtype=qt.dtype)
        b = torch.zeros((0,), dtype=qt.dtype)
    ts = [v if isinstance(v, torch.Tensor) else torch.as_tensor(v) for v in (Q, q, G, h)]
    At = A if isinstance(A, torch.Tensor) else torch.as_tensor(A)
    bt_ = b if isinstance(b, torch.Tensor) else torch.as_tensor(b)
    return _SolveQPElasticPrimal.apply(*ts, penalty, At, bt_, solver_tol, target_kappa, mode, relax, exact_dG)


def initialize_elastic(Q, q, G, h, penalty, A=None, b=None):
    """qpax.elastic_qp.initialize_elastic (elastic_qp.py:27-47) -> (x, t, s1, s2, z1, z2), with Q
    symmetrised as solve_qp_elastic hands it over (:186); y is appended with equality rows."""
    bt, pen = _prep(Q, q, G, h, penalty, A, b)
    lib = _lib.load()
    n, me, mi = bt.n, bt.me, bt.mi
    outs = [bt.out(n), bt.out(mi), bt.out(mi), bt.out(mi), bt.out(mi), bt.out(mi), bt.out(me)]
    ws, nbytes = _workspace(bt)
    st = lib.qpx_initialize_elastic(dtype_code(bt.dtype), bt.B, n, me, mi, *[ptr(bt[k]) for k in bt.names], ptr(pen),
                                    *[ptr(o) for o in outs], ptr(ws), nbytes, stream_of(bt.device))
    _lib.check(st, "qpx_initialize_elastic")
    res = [bt.unflatten(outs[0], n)] + [bt.unflatten(v, mi) for v in outs[1:6]]
    if A is not None:
        res.append(bt.unflatten(outs[6], me))
    return tuple(res)
