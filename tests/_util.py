"""Shared helpers for the parity tests."""

import numpy as np


def rel_err(a, b):
    """Per-problem relative 2-norm error ||a-b|| / max(||b||, 1e-12) over all but the batch axis.
    Entries that are infinite in both (s on rows with infinite h) count as equal; an infinity in
    one only makes the error infinite."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    both = np.isinf(a) & np.isinf(b) & (np.sign(a) == np.sign(b))
    one = (np.isinf(a) | np.isinf(b)) & ~both
    a = np.where(both | one, 0.0, a)
    b = np.where(both | one, 0.0, b)
    err = np.linalg.norm((a - b).reshape(a.shape[0], -1), axis=1) / np.maximum(
        np.linalg.norm(b.reshape(b.shape[0], -1), axis=1), 1e-12)
    return np.where(one.reshape(one.shape[0], -1).any(axis=1), np.inf, err)


def to_np(t):
    return t.detach().cpu().numpy()


def conditioning_floor(solve, data, ref, ok, reps=2, seed=0):
    """Per output (x, s, z, y): the largest relative change of the float64 oracle's own result,
    over the problems in `ok`, when every input entry is scaled by (1 +- 2^-52) -- how far two
    correct float64 solvers that round differently can end up apart on these problems.  `solve`
    maps the six inputs to (x, s, z, y, converged, iters); `ref` is its result on `data`."""
    rng = np.random.default_rng(seed)
    floor = {}
    for _ in range(reps):
        dp = [v * (1.0 + 2.220446049250313e-16 * rng.choice([-1.0, 1.0], size=v.shape)) for v in data]
        out = solve(*dp)
        same = ok & (out[4] == ref[4]) & (out[5] == ref[5])
        for i, k in enumerate("xszy"):
            if ref[i].shape[-1] and same.any():
                floor[k] = max(floor.get(k, 0.0), float(rel_err(out[i][same], ref[i][same]).max()))
    return floor
