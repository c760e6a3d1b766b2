"""Argument normalisation shared by the qpax-compatible entry points.

qpax takes one problem per call and is batched with ``jax.vmap``; here the same functions take
torch tensors with any number of leading batch axes (broadcast against each other, as ``vmap``
with ``in_axes=None`` would), which is what ``vmap`` produces.  ``Q``, ``A`` and ``G`` follow
``jnp.atleast_2d`` (qpax/pdip.py:204-206): the row count of ``A``/``G`` is taken from ``b``/``h``.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib


def as_tensor(v, device=None):
    if isinstance(v, torch.Tensor):
        return v
    return torch.as_tensor(np.asarray(v))


def pick_device(*ts):
    for t in ts:
        if isinstance(t, torch.Tensor) and t.is_cuda:
            return t.device
    if not torch.cuda.is_available():
        raise RuntimeError("qpax_amd needs a ROCm GPU: torch.cuda.is_available() is False")
    return torch.device("cuda", torch.cuda.current_device())


def result_dtype(*ts):
    dt = torch.float32
    for t in ts:
        if t.dtype == torch.float64:
            dt = torch.float64
    return dt


def dtype_code(dt):
    if dt == torch.float32:
        return _lib.QPX_F32
    if dt == torch.float64:
        return _lib.QPX_F64
    raise TypeError(f"unsupported dtype {dt}")


class Batch:
    """Flattened, contiguous (B, ...) views of one call's operands on one device."""

    def __init__(self, names, tensors, batch_shape, n, me, mi, dtype, device):
        self.names = names
        self.t = dict(zip(names, tensors))
        self.batch_shape = batch_shape
        self.B = int(np.prod(batch_shape)) if len(batch_shape) else 1
        self.n, self.me, self.mi = n, me, mi
        self.dtype = dtype
        self.device = device

    def __getitem__(self, k):
        return self.t[k]

    def out(self, *tail, dtype=None):
        return torch.empty((self.B, *tail), dtype=dtype or self.dtype, device=self.device)

    def unflatten(self, t, *tail):
        return t.reshape(tuple(self.batch_shape) + tuple(tail))


def normalize(Q, q, A, b, G, h, device=None):
    """Returns a Batch of (Q, q, A, b, G, h) with shapes (B,n,n) (B,n) (B,me,n) (B,me) (B,mi,n)
    (B,mi).  ``device`` defaults to the inputs' GPU (or the current one)."""
    raw = [as_tensor(v) for v in (Q, q, A, b, G, h)]
    device = device if device is not None else pick_device(*raw)
    dt = result_dtype(*raw)
    Q, q, A, b, G, h = [t.to(device=device, dtype=dt) for t in raw]
    n = q.shape[-1]
    me = b.shape[-1] if b.dim() >= 1 else 0
    mi = h.shape[-1] if h.dim() >= 1 else 0
    if Q.dim() < 2:
        Q = Q.reshape(*Q.shape[:-1], 1, -1) if Q.dim() == 1 else Q.reshape(1, 1)
    bs = [Q.shape[:-2], q.shape[:-1], b.shape[:-1], h.shape[:-1]]
    A = _rows(A, me, n)
    G = _rows(G, mi, n)
    bs += [A.shape[:-2], G.shape[:-2]]
    batch = torch.broadcast_shapes(*bs)
    Q = Q.expand(*batch, n, n)
    q = q.expand(*batch, n)
    A = A.expand(*batch, me, n)
    b = b.expand(*batch, me)
    G = G.expand(*batch, mi, n)
    h = h.expand(*batch, mi)
    B = int(np.prod(batch)) if len(batch) else 1
    ts = [t.reshape(B, *t.shape[len(batch):]).contiguous() for t in (Q, q, A, b, G, h)]
    return Batch(("Q", "q", "A", "b", "G", "h"), ts, tuple(batch), n, me, mi, dt, device)


def _rows(M, m, n):
    """atleast_2d plus the row count known from the right-hand side vector."""
    if M.dim() == 1:
        return M.reshape(m, n) if M.numel() == m * n else M.reshape(1, -1)
    return M.reshape(*M.shape[:-2], m, n)


def ptr(t):
    return ctypes_ptr(t)


def ctypes_ptr(t):
    return t.data_ptr() if t is not None and t.numel() > 0 else None


def stream_of(device):
    return torch.cuda.current_stream(device).cuda_stream
