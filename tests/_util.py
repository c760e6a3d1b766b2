"""Shared helpers for the parity tests."""

import numpy as np


def rel_err(a, b):
    """Per-problem relative 2-norm error ||a-b|| / max(||b||, 1e-12) over all but the batch axis."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    ax = tuple(range(1, a.ndim))
    return np.linalg.norm((a - b).reshape(a.shape[0], -1), axis=1) / np.maximum(
        np.linalg.norm(b.reshape(b.shape[0], -1), axis=1), 1e-12)


def to_np(t):
    return t.detach().cpu().numpy()
