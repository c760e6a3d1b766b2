"""Host-side argument handling (no GPU): jnp.atleast_2d semantics, broadcasting of batch axes
(the reference's vmap), dtype promotion."""

import numpy as np
import pytest
import torch

from qpax_amd._tensors import normalize


def test_unbatched_shapes():
    bt = normalize(np.eye(3), np.ones(3), np.array([1.0, 1.0, 1.0]), np.array([1.0]), np.zeros((0, 3)), np.zeros(0),
                   device="cpu")
    assert (bt.B, bt.n, bt.me, bt.mi) == (1, 3, 1, 0)
    assert bt["A"].shape == (1, 1, 3)  # atleast_2d of a 1-D A (pdip.py:205)
    assert bt.unflatten(torch.zeros(1)).shape == ()
    assert bt.unflatten(torch.zeros(1, 3), 3).shape == (3,)


def test_broadcast_batch_axes():
    B = 4
    Q = np.broadcast_to(np.eye(2), (B, 2, 2))
    q = np.ones((B, 2))
    G = -np.eye(2)  # shared across the batch (vmap in_axes=None)
    h = np.zeros(2)
    bt = normalize(Q, q, np.zeros((0, 2)), np.zeros(0), G, h, device="cpu")
    assert bt.B == B and bt["G"].shape == (B, 2, 2) and bt["h"].shape == (B, 2)
    assert bt["A"].shape == (B, 0, 2)


def test_nested_batch_axes():
    Q = np.broadcast_to(np.eye(2), (2, 3, 2, 2))
    bt = normalize(Q, np.ones((2, 3, 2)), np.zeros((0, 2)), np.zeros(0), -np.eye(2), np.zeros(2), device="cpu")
    assert bt.batch_shape == (2, 3) and bt.B == 6
    assert bt.unflatten(torch.zeros(6, 2), 2).shape == (2, 3, 2)


@pytest.mark.parametrize("dts,expect", [((np.float32,) * 6, torch.float32),
                                        ((np.float32,) * 5 + (np.float64,), torch.float64)])
def test_dtype_promotion(dts, expect):
    args = [np.eye(2), np.ones(2), np.zeros((0, 2)), np.zeros(0), -np.eye(2), np.zeros(2)]
    bt = normalize(*[np.asarray(a, d) for a, d in zip(args, dts)], device="cpu")
    assert bt.dtype == expect
    assert all(bt[k].dtype == expect for k in bt.names)


def test_no_gpu_raises_cleanly():
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    import qpax_amd

    with pytest.raises(RuntimeError, match="GPU"):
        qpax_amd.solve_qp(np.eye(2), np.ones(2), np.zeros((0, 2)), np.zeros(0), -np.eye(2), np.zeros(2))
