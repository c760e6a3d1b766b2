"""The static shape branches of qpax.solve_qp for problems without inequality rows
(qpax/pdip.py:208-217), on the HIP library: no constraints -> x = solve(Q, -q); equality rows
only -> solve_eq_only (pdip.py:137-150)."""

from __future__ import annotations

import torch

from . import _lib
from ._tensors import dtype_code, ptr, stream_of


def _run(bt):
    lib = _lib.load()
    x, y = bt.out(bt.n), bt.out(bt.me)
    flags = torch.empty(2, bt.B, dtype=torch.int32, device=bt.device)
    nbytes = lib.qpx_special_workspace_size(bt.B, bt.n, bt.me)
    ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=bt.device)
    st = lib.qpx_solve_qp_special(dtype_code(bt.dtype), bt.B, bt.n, bt.me, ptr(bt["Q"]), ptr(bt["q"]), ptr(bt["A"]),
                                  ptr(bt["b"]), ptr(x), ptr(y), ptr(flags[0]), ptr(flags[1]), ptr(ws), nbytes,
                                  stream_of(bt.device))
    _lib.check(st, "qpx_solve_qp_special")
    return x, y


def solve_unconstrained(bt):
    return _run(bt)[0]


def solve_eq_only(bt):
    return _run(bt)
