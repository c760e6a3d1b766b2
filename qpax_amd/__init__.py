"""qpax_amd -- qpax's batched, differentiable QP solver on AMD MI355X (gfx950).

Drop-in entry points with qpax's names, arguments and return tuples (qpax/__init__.py:1-3):

* ``solve_qp(Q, q, A, b, G, h, solver_tol=1e-3) -> (x, s, z, y, converged, iters)``
* ``solve_qp_primal(Q, q, A, b, G, h, solver_tol=1e-5, target_kappa=1e-3) -> x`` (differentiable)
* ``solve_qp_elastic(Q, q, G, h, penalty, solver_tol=1e-3) -> (x, t, s1, s2, z1, z2, converged, iters)``
* ``solve_qp_elastic_primal(Q, q, G, h, penalty, solver_tol=1e-5, target_kappa=1e-3) -> x``

Inputs are torch tensors with optional leading batch axes (the reference's ``jax.vmap``); all
arithmetic runs in the HIP library ``libqpx.so`` (C ABI in include/qpx.h).
"""

from .diff_qp import solve_qp_primal
from .elastic_qp import solve_qp_elastic, solve_qp_elastic_primal
from .pdip import solve_qp

__all__ = ["solve_qp", "solve_qp_primal", "solve_qp_elastic", "solve_qp_elastic_primal"]
