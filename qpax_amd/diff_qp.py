"""qpax.solve_qp_primal with reverse-mode differentiation -- host side of qpax/diff_qp.py.

``solve_qp_primal(Q, q, A, b, G, h, solver_tol=1e-5, target_kappa=1e-3) -> x`` is a
``torch.autograd.Function``: its forward runs the library's solve (and, when any input requires
grad, the kappa-relaxation of qpax/pdip_relaxed.py), its backward runs the library's implicit
differentiation (diff_qp / optnet_derivatives, qpax/diff_qp.py:8-46) and returns cotangents for
the six arrays and None for the two floats, like the reference's custom_vjp (diff_qp.py:74-82).
"""

from __future__ import annotations

import torch

from . import _lib
from ._tensors import dtype_code, normalize, ptr, stream_of
from .pdip import _workspace, solve_qp_batch, solve_qp_relaxed_batch


def diff_qp_batch(bt, relaxed, dl_dx, mode="fused", exact_dG=False):
    """Six cotangents (flat) at the relaxed point ``relaxed = (xr, sr, zr, yr)``."""
    lib = _lib.load()
    xr, sr, zr, yr = relaxed
    g = dl_dx.to(bt.dtype).reshape(bt.B, bt.n).contiguous()
    dQ, dq, dA = bt.out(bt.n, bt.n), bt.out(bt.n), bt.out(bt.me, bt.n)
    db, dG, dh = bt.out(bt.me), bt.out(bt.mi, bt.n), bt.out(bt.mi)
    ws, nbytes = _workspace(bt)
    st = lib.qpx_diff_qp(dtype_code(bt.dtype), _lib.MODES[mode], bt.B, bt.n, bt.me, bt.mi,
                         *[ptr(bt[k]) for k in bt.names], ptr(xr), ptr(sr), ptr(zr), ptr(yr), ptr(g),
                         _lib.QPX_EXACT_DG if exact_dG else 0, ptr(dQ), ptr(dq), ptr(dA), ptr(db), ptr(dG), ptr(dh), ptr(ws), nbytes,
                         stream_of(bt.device))
    _lib.check(st, "qpx_diff_qp")
    return dQ, dq, dA, db, dG, dh


def _reduce_to(grad, shape):
    """Sum a broadcast cotangent back to the operand's own shape."""
    if tuple(grad.shape) == tuple(shape):
        return grad
    while grad.dim() > len(shape):
        grad = grad.sum(0)
    for i, (gs, s) in enumerate(zip(grad.shape, shape)):
        if s == 1 and gs != 1:
            grad = grad.sum(i, keepdim=True)
    return grad.reshape(shape)


class _SolveQPPrimal(torch.autograd.Function):
    @staticmethod
    def forward(ctx, Q, q, A, b, G, h, solver_tol, target_kappa, mode, exact_dG):
        bt = normalize(Q, q, A, b, G, h)
        if bt.mi == 0:
            from .pdip import _special

            return _special(bt)[0]
        need = any(ctx.needs_input_grad[:6])
        if need:
            # no read-back of the relaxation flags here (it would hold the host until the forward
            # ends): a relaxation that did not start leaves NaN at its relaxed point, so the
            # backward returns NaN cotangents for that problem
            out = solve_qp_relaxed_batch(bt, solver_tol, target_kappa, mode, check=False)
            x, relaxed = out[0], out[6:10]
            ctx.save_for_backward(*[bt[k] for k in bt.names], *relaxed)
            ctx.meta = (bt.batch_shape, bt.n, bt.me, bt.mi, mode, [tuple(t.shape) for t in (Q, q, A, b, G, h)],
                        exact_dG)
        else:
            x = solve_qp_batch(bt, solver_tol, mode)[0]
        return bt.unflatten(x, bt.n)

    @staticmethod
    def backward(ctx, gx):
        Q, q, A, b, G, h, xr, sr, zr, yr = ctx.saved_tensors
        batch_shape, n, me, mi, mode, shapes, exact_dG = ctx.meta
        bt = normalize(Q, q, A, b, G, h)
        grads = diff_qp_batch(bt, (xr, sr, zr, yr), gx.contiguous(), mode, exact_dG)
        tails = [(n, n), (n,), (me, n), (me,), (mi, n), (mi,)]
        out = []
        for gr, tail, shp, need in zip(grads, tails, shapes, ctx.needs_input_grad[:6]):
            out.append(_reduce_to(gr.reshape(*batch_shape, *tail), shp) if need else None)
        return (*out, None, None, None, None)


def solve_qp_primal(Q, q, A, b, G, h, solver_tol: float = 1e-5, target_kappa: float = 1e-3, mode: str = "fused",
                    exact_dG: bool = False):
    """Primal solution x of the QP, differentiable w.r.t. (Q, q, A, b, G, h) (qpax/diff_qp.py:49).

    ``exact_dG=True`` returns the exact vector-Jacobian product for G; the default reproduces the
    reference's formula (qpax/diff_qp.py:11), which scales the second term of dG by lam."""
    ts = [v if isinstance(v, torch.Tensor) else torch.as_tensor(v) for v in (Q, q, A, b, G, h)]
    return _SolveQPPrimal.apply(*ts, solver_tol, target_kappa, mode, exact_dG)
