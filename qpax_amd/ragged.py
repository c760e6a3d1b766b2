"""Ragged batches: QPs of different sizes solved and differentiated in one library call.

qpax batches one problem size with ``jax.vmap``; a batch of mixed sizes there is a Python loop
of per-size calls.  Here a :class:`RaggedPlan` (the C-ABI plan of ``qpx_plan_create``) takes the
sizes once, and every call hands the library the operands packed back to back: one launch per
kernel bucket, with the problems of a bucket pulled from a device work queue by persistent
workgroups, largest first.  Per problem the arithmetic is that of ``solve_qp`` /
``solve_qp_primal`` in fused mode (qpax/pdip.py:181-251, qpax/diff_qp.py:50-79).

Calls allocate nothing inside the library and never synchronise, so a whole forward + backward
can be captured in a ``torch.cuda.CUDAGraph`` (see :meth:`RaggedPlan.capture`).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._tensors import dtype_code, stream_of

_NAMES = ("Q", "q", "A", "b", "G", "h")


def _p(t):
    return t.data_ptr() if t is not None and t.numel() > 0 else None


class RaggedPlan:
    """Schedule of a ragged batch on the current GPU.

    ``n``, ``m_eq``, ``m_ineq``: per-problem sizes (length-B integer sequences).  The plan owns
    its device tables and a workspace; calls through one plan are ordered on the current stream.
    """

    def __init__(self, n, m_eq, m_ineq, dtype=torch.float32, device=None):
        self.n = np.ascontiguousarray(n, dtype=np.int32)
        self.me = np.ascontiguousarray(m_eq, dtype=np.int32)
        self.mi = np.ascontiguousarray(m_ineq, dtype=np.int32)
        if not (self.n.shape == self.me.shape == self.mi.shape) or self.n.ndim != 1:
            raise ValueError("n, m_eq, m_ineq must be 1-D sequences of one length")
        if not torch.cuda.is_available():
            raise RuntimeError("qpax_amd needs a ROCm GPU: torch.cuda.is_available() is False")
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.dtype = dtype
        self.B = int(self.n.shape[0])
        lib = _lib.load()
        self._lib = lib
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(lib.qpx_plan_create(dtype_code(dtype), self.B, self.n.ctypes.data, self.me.ctypes.data,
                                           self.mi.ctypes.data, ctypes.byref(handle)), "qpx_plan_create")
        self._h = handle
        info = np.zeros(8, dtype=np.int64)
        _lib.check(lib.qpx_plan_info(self._h, info.ctypes.data), "qpx_plan_info")
        self.totals = dict(zip(_NAMES, (int(v) for v in info[:6])))
        self.ws_bytes = int(info[6])
        self.n_buckets = int(info[7])
        off = np.zeros((max(self.B, 1), 6), dtype=np.int64)
        _lib.check(lib.qpx_plan_offsets(self._h, off.ctypes.data), "qpx_plan_offsets")
        self.offsets = off[: self.B]
        self.workspace = torch.empty(max(self.ws_bytes, 1), dtype=torch.uint8, device=self.device)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.qpx_plan_destroy(h)
            self._h = None

    # ------------------------------------------------------------------ packing
    def shapes(self, i):
        n, me, mi = int(self.n[i]), int(self.me[i]), int(self.mi[i])
        return {"Q": (n, n), "q": (n,), "A": (me, n), "b": (me,), "G": (mi, n), "h": (mi,)}

    def pack(self, problems):
        """Packs a list of (Q, q, A, b, G, h) per-problem arrays into six flat device tensors."""
        if len(problems) != self.B:
            raise ValueError(f"expected {self.B} problems, got {len(problems)}")
        host = {k: np.empty(self.totals[k], dtype=np.float64) for k in _NAMES}
        for i, pr in enumerate(problems):
            sh = self.shapes(i)
            for j, k in enumerate(_NAMES):
                a = np.asarray(pr[j], dtype=np.float64).reshape(sh[k])
                o = self.offsets[i, j]
                host[k][o:o + a.size] = a.ravel()
        return tuple(torch.tensor(host[k], dtype=self.dtype, device=self.device) for k in _NAMES)

    def unpack(self, flat, which, i):
        """Problem i's slice of a packed tensor laid out like operand ``which`` ("Q".."h")."""
        j = _NAMES.index(which)
        sh = self.shapes(i)[which]
        o = int(self.offsets[i, j])
        return flat[o:o + int(np.prod(sh))].reshape(sh)

    def empty_like_operand(self, which, dtype=None):
        return torch.empty(max(self.totals[which], 0), dtype=dtype or self.dtype, device=self.device)

    def outputs(self, relax=True):
        """Fresh output buffers: x, s, z, y, converged, iters[, xr, sr, zr, yr, rconv, riters]."""
        e = self.empty_like_operand
        i32 = dict(dtype=torch.int32, device=self.device)
        out = [e("q"), e("h"), e("h"), e("b"), torch.empty(self.B, **i32), torch.empty(self.B, **i32)]
        if relax:
            out += [e("q"), e("h"), e("h"), e("b"), torch.empty(self.B, **i32), torch.empty(self.B, **i32)]
        return out

    def grad_outputs(self):
        e = self.empty_like_operand
        return [e(k) for k in _NAMES]

    def _check(self, ts):
        for k, t in zip(_NAMES, ts):
            if t.dtype != self.dtype or t.device != self.device or not t.is_contiguous() or t.numel() != self.totals[k]:
                raise ValueError(f"packed {k} must be a contiguous {self.dtype} tensor of {self.totals[k]} elements "
                                 f"on {self.device}")

    # ------------------------------------------------------------------ calls
    def solve(self, ops, solver_tol=1e-3, out=None):
        """solve_qp over the ragged batch; returns packed (x, s, z, y, converged, iters)."""
        self._check(ops)
        out = out if out is not None else self.outputs(relax=False)
        st = self._lib.qpx_plan_solve(self._h, *[_p(t) for t in ops], float(solver_tol), *[_p(t) for t in out[:6]],
                                      _p(self.workspace), self.ws_bytes, stream_of(self.device))
        _lib.check(st, "qpx_plan_solve")
        return tuple(out[:6])

    def new_workspace(self):
        """A workspace of this plan's size, for calls that run at the same time as others on the
        plan (each concurrent call needs its own)."""
        return torch.empty(max(self.ws_bytes, 1), dtype=torch.uint8, device=self.device)

    def solve_relaxed(self, ops, solver_tol=1e-5, target_kappa=1e-3, out=None, tune=False, workspace=None):
        """solve_qp then relax_qp (the forward of solve_qp_primal); returns the 12 packed outputs."""
        self._check(ops)
        out = out if out is not None else self.outputs(relax=True)
        ws = workspace if workspace is not None else self.workspace
        fn = self._lib.qpx_plan_tune if tune else self._lib.qpx_plan_solve_relaxed
        st = fn(self._h, *[_p(t) for t in ops], float(solver_tol), float(target_kappa), *[_p(t) for t in out],
                _p(ws), self.ws_bytes, stream_of(self.device))
        _lib.check(st, "qpx_plan_tune" if tune else "qpx_plan_solve_relaxed")
        return tuple(out)

    def tune(self, ops, solver_tol=1e-5, target_kappa=1e-3):
        """Autotunes every bucket's launch shape on these operands (synchronises)."""
        return self.solve_relaxed(ops, solver_tol, target_kappa, tune=True)

    def diff(self, ops, relaxed, dl_dx, exact_dG=False, out=None, workspace=None):
        """Backward of solve_qp_primal at the relaxed point (xr, sr, zr, yr); returns packed
        (dQ, dq, dA, db, dG, dh)."""
        self._check(ops)
        out = out if out is not None else self.grad_outputs()
        ws = workspace if workspace is not None else self.workspace
        flags = _lib.QPX_EXACT_DG if exact_dG else 0
        st = self._lib.qpx_plan_diff(self._h, *[_p(t) for t in ops], *[_p(t) for t in relaxed], _p(dl_dx), flags,
                                     *[_p(t) for t in out], _p(ws), self.ws_bytes,
                                     stream_of(self.device))
        _lib.check(st, "qpx_plan_diff")
        return tuple(out)

    def buckets(self):
        """Per-bucket schedule: list of dicts (kernel kind, instance, launch shape, problems)."""
        rows = []
        for i in range(self.n_buckets):
            o = np.zeros(10, dtype=np.int64)
            _lib.check(self._lib.qpx_plan_bucket(self._h, i, o.ctypes.data), "qpx_plan_bucket")
            kind = {0: "wave", 1: "block", 2: "group"}[int(o[0])]
            rows.append(dict(kind=kind, nmax=int(o[1]), mic=int(o[2]), threads=int(o[3]),
                             waves_per_group=int(o[4]), persistent=bool(o[5]), problems=int(o[6]),
                             lds_per_wave=int(o[7]), groups=int(o[8]), flops_per_iter=int(o[9])))
        return rows

    def capture(self, ops, dl_dx, solver_tol=1e-5, target_kappa=1e-3, exact_dG=False):
        """Captures forward (solve + relax) and backward into one CUDA graph over static buffers.

        Returns (graph, fwd_outputs, grads): ``graph.replay()`` recomputes both from the current
        contents of ``ops`` and ``dl_dx``."""
        fwd = self.outputs(relax=True)
        grads = self.grad_outputs()
        relaxed = (fwd[6], fwd[7], fwd[8], fwd[9])
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):  # warm-up outside the capture
            self.solve_relaxed(ops, solver_tol, target_kappa, out=fwd)
            self.diff(ops, relaxed, dl_dx, exact_dG, out=grads)
        torch.cuda.current_stream(self.device).wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            self.solve_relaxed(ops, solver_tol, target_kappa, out=fwd)
            self.diff(ops, relaxed, dl_dx, exact_dG, out=grads)
        return g, fwd, grads


class _RaggedPrimal(torch.autograd.Function):
    @staticmethod
    def forward(ctx, plan, solver_tol, target_kappa, exact_dG, Q, q, A, b, G, h):
        ops = (Q.contiguous(), q.contiguous(), A.contiguous(), b.contiguous(), G.contiguous(), h.contiguous())
        out = plan.solve_relaxed(ops, solver_tol, target_kappa)
        ctx.plan, ctx.exact_dG = plan, exact_dG
        ctx.save_for_backward(*ops, out[6], out[7], out[8], out[9])
        return out[0]

    @staticmethod
    def backward(ctx, gx):
        Q, q, A, b, G, h, xr, sr, zr, yr = ctx.saved_tensors
        grads = ctx.plan.diff((Q, q, A, b, G, h), (xr, sr, zr, yr), gx.contiguous(), ctx.exact_dG)
        return (None, None, None, None) + tuple(grads)


def solve_qp_primal_ragged(plan: RaggedPlan, Q, q, A, b, G, h, solver_tol=1e-5, target_kappa=1e-3, exact_dG=False):
    """solve_qp_primal (qpax/diff_qp.py:50-79) over a ragged batch of packed operands: returns the
    packed primal solutions; differentiable w.r.t. all six packed operands."""
    return _RaggedPrimal.apply(plan, solver_tol, target_kappa, exact_dG, Q, q, A, b, G, h)
