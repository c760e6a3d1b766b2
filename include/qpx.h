/* libqpx -- batched dense convex QP solve (primal-dual interior point) and its implicit-
 * differentiation backward on AMD Instinct MI355X (gfx950).
 *
 * C ABI: plain pointers and sizes only.  All array arguments are DEVICE pointers to row-major,
 * contiguous arrays with a leading batch axis, in the element type named by `dtype`
 * (QPX_F32 = float, QPX_F64 = double); flags are int32.  `stream` is a hipStream_t (NULL = the
 * default stream).  Every call is stream-ordered; the caller owns every buffer, including the
 * workspace, and the library keeps no pointer past the call.
 *
 * Problem (per batch entry):   minimise 1/2 x'Qx + q'x   s.t.  A x = b,  G x <= h
 *   Q (n,n)  q (n)  A (m_eq,n)  b (m_eq)  G (m_ineq,n)  h (m_ineq)
 *
 * Numerical failure is not an error: it is reported per problem through converged = 0, as the
 * reference does.  Errors (bad arguments, HIP failures) return a non-zero status and leave a
 * message for qpx_last_error().
 *
 * Modes: QPX_PLAIN runs one kernel per IPM stage per iteration and checks convergence on the host
 * (one device->host read per iteration); QPX_FUSED runs the whole loop in one kernel launch per
 * phase with no host synchronisation (capturable in a hipGraph).  The two modes run different
 * eliminations for the problems the fused kernels take, so their results agree to rounding (x within
 * about 1e-9 relative in float64 on the test problems), not bit for bit.
 */
#ifndef QPX_H_
#define QPX_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QPX_API __attribute__((visibility("default")))

typedef enum qpx_status {
  QPX_OK = 0,
  QPX_ERR_ARG = 1,          /* invalid argument (sizes, null pointers, dtype, mode) */
  QPX_ERR_HIP = 2,          /* a HIP runtime call failed */
  QPX_ERR_WORKSPACE = 3,    /* workspace smaller than qpx_workspace_size() */
  QPX_ERR_UNSUPPORTED = 4   /* a size outside what this build supports */
} qpx_status;

enum { QPX_F32 = 0, QPX_F64 = 1 };
enum { QPX_PLAIN = 0, QPX_FUSED = 1 };
/* flags of the backward and relaxation entry points */
enum {
  QPX_EXACT_DG = 1,           /* dG = diag(lam) dlam x^T + lam dx^T, the exact VJP (default: the
                                 reference's diag(lam) (dlam x^T + lam dx^T), qpax/diff_qp.py:11) */
  QPX_ELASTIC_CONSISTENT = 2, /* elastic relaxation with the forward's signs of t and r2 (default:
                                 the reference's flipped signs, qpax/elastic_qp.py:257-261) */
  QPX_TRACE_DEBUG_STEP = 4    /* qpx_solve_qp_trace: the step length rule of the reference's debug
                                 loop, 0.99 min(1, 0.99 ort(s), 0.99 ort(z)) (qpax/pdip.py:281) */
};

/* Library version (major*10000 + minor*100 + patch). */
QPX_API int32_t qpx_version(void);

/* Thread-local text of the last error returned on this thread ("" if none). */
QPX_API const char* qpx_last_error(void);

/* Bytes of device workspace a call with these sizes needs (any entry point below). */
QPX_API size_t qpx_workspace_size(int64_t batch, int32_t n, int32_t m_eq, int32_t m_ineq);

/* qpax.solve_qp -- replaces qpax/pdip.py:181-251 (solve_qp), the driver of initialize
 * (pdip.py:15-35), the predictor-corrector loop (pdip.py:87-119, :228-238), inf-bound removal
 * (pdip.py:166-178, :243-246) and Q symmetrisation (pdip.py:223).  Requires m_ineq >= 1 (the
 * reference's static branches for m_ineq == 0 are qpx_solve_qp_special).
 * Outputs: x (B,n), s (B,m_ineq), z (B,m_ineq), y (B,m_eq), converged (B), iters (B). */
QPX_API qpx_status qpx_solve_qp(int32_t dtype, int32_t mode, int64_t batch, int32_t n,
                                int32_t m_eq, int32_t m_ineq, const void* Q, const void* q,
                                const void* A, const void* b, const void* G, const void* h,
                                double solver_tol, void* x, void* s, void* z, void* y,
                                int32_t* converged, int32_t* iters, void* workspace,
                                size_t workspace_bytes, void* stream);

/* Forward of qpax.solve_qp_primal under differentiation -- replaces
 * solve_qp_primal_forward (qpax/diff_qp.py:61-71): solve_qp, then relax_qp
 * (qpax/pdip_relaxed.py:70-103, up to 30 damped Newton steps on s*z = target_kappa) from the
 * solution.  Writes the solve outputs as qpx_solve_qp plus the relaxed point xr, sr, zr, yr and
 * the relaxation's converged/iters flags. */
QPX_API qpx_status qpx_solve_qp_relaxed(int32_t dtype, int32_t mode, int64_t batch, int32_t n,
                                        int32_t m_eq, int32_t m_ineq, const void* Q, const void* q,
                                        const void* A, const void* b, const void* G, const void* h,
                                        double solver_tol, double target_kappa, void* x, void* s,
                                        void* z, void* y, int32_t* converged, int32_t* iters,
                                        void* xr, void* sr, void* zr, void* yr,
                                        int32_t* relax_converged, int32_t* relax_iters,
                                        void* workspace, size_t workspace_bytes, void* stream);

/* qpx_solve_qp_relaxed (QPX_FUSED) without a workspace argument and without any allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc).  Its scratch is a
 * device array of qpx_noalloc_scratch_bytes() bytes (256 MiB) inside the library, one per GPU,
 * held until the process ends; a call needs 4 KiB plus, per problem, 8 * (round64(m^2) + 640) bytes
 * for sizes the wavefront kernels take (m = n rounded up to a multiple of 4, plus m_eq; 36 KB at
 * n = 50, m_eq = 10, so 4096 such problems fit) and
 * qpx_workspace_size(1, ...) - 4096 bytes for the others, and returns QPX_ERR_UNSUPPORTED beyond
 * the array.  Calls on one GPU share the array: calls that overlap on one GPU (two streams, or two
 * host threads) race.  Results are those of qpx_solve_qp_relaxed in QPX_FUSED mode, bit for bit. */
QPX_API qpx_status qpx_solve_qp_relaxed_noalloc(int32_t dtype, int64_t batch, int32_t n, int32_t m_eq,
                                                int32_t m_ineq, const void* Q, const void* q, const void* A,
                                                const void* b, const void* G, const void* h, double solver_tol,
                                                double target_kappa, void* x, void* s, void* z, void* y,
                                                int32_t* converged, int32_t* iters, void* xr, void* sr, void* zr,
                                                void* yr, int32_t* relax_converged, int32_t* relax_iters,
                                                void* stream);
/* Bytes of the no-allocation entry point's scratch array. */
QPX_API size_t qpx_noalloc_scratch_bytes(void);

/* Backward of qpax.solve_qp_primal -- replaces solve_qp_primal_backward / diff_qp /
 * optnet_derivatives (qpax/diff_qp.py:8-46, :74-79): factor the KKT system at the relaxed point
 * (sr, zr), solve it with right-hand side (-dl_dx, 0, 0, 0) and form the six cotangents
 * dQ (B,n,n), dq (B,n), dA (B,m_eq,n), db (B,m_eq), dG (B,m_ineq,n), dh (B,m_ineq).
 * flags: 0 or QPX_EXACT_DG. */
QPX_API qpx_status qpx_diff_qp(int32_t dtype, int32_t mode, int64_t batch, int32_t n,
                               int32_t m_eq, int32_t m_ineq, const void* Q, const void* q,
                               const void* A, const void* b, const void* G, const void* h,
                               const void* xr, const void* sr, const void* zr, const void* yr,
                               const void* dl_dx, int32_t flags, void* dQ, void* dq, void* dA,
                               void* db, void* dG, void* dh, void* workspace,
                               size_t workspace_bytes, void* stream);

/* The static shape branches of qpax.solve_qp for m_ineq == 0 (qpax/pdip.py:208-217): with
 * m_eq == 0, x = solve(Q, -q) (pdip.py:210); with m_eq > 0, solve_eq_only (pdip.py:137-150).  Both
 * use the unsymmetrised Q, factored by LU with partial pivoting.  converged = 1, iters = 0.
 * Workspace: qpx_special_workspace_size(batch, n, m_eq) bytes. */
QPX_API qpx_status qpx_solve_qp_special(int32_t dtype, int64_t batch, int32_t n, int32_t m_eq,
                                        const void* Q, const void* q, const void* A, const void* b,
                                        void* x, void* y, int32_t* converged, int32_t* iters,
                                        void* workspace, size_t workspace_bytes, void* stream);
QPX_API size_t qpx_special_workspace_size(int64_t batch, int32_t n, int32_t m_eq);

/* qpax.pdip.initialize -- replaces qpax/pdip.py:15-35 (the CVXOPT initialisation: one reduced KKT
 * solve with D = 1 and the shift of s, z into the positive orthant), on the problem as solve_qp
 * hands it over: Q symmetrised, rows with infinite h zeroed (pdip.py:220-223).  Outputs x (B,n),
 * s, z (B,m_ineq), y (B,m_eq).  Workspace: qpx_workspace_size(batch, n, m_eq, m_ineq). */
QPX_API qpx_status qpx_initialize(int32_t dtype, int64_t batch, int32_t n, int32_t m_eq, int32_t m_ineq,
                                  const void* Q, const void* q, const void* A, const void* b, const void* G,
                                  const void* h, void* x, void* s, void* z, void* y, void* workspace,
                                  size_t workspace_bytes, void* stream);

/* qpax.solve_qp with a per-iteration trace -- the device-side counterpart of solve_qp_debug
 * (qpax/pdip.py:254-360), which prints one line per predictor-corrector iteration.  Solves as
 * qpx_solve_qp (the workgroup-per-QP kernels) and writes trace (B, 30, 5) float64: for iteration k,
 * the infinity norms of r1, r2, r3, r4 at the iteration's entry (r4: 0 without equality rows) and
 * the step length taken; rows of iterations not taken are NaN.  flags: 0 (the step rule of
 * solve_qp, pdip.py:112) or QPX_TRACE_DEBUG_STEP (the debug loop's rule, pdip.py:281). */
QPX_API qpx_status qpx_solve_qp_trace(int32_t dtype, int64_t batch, int32_t n, int32_t m_eq, int32_t m_ineq,
                                      const void* Q, const void* q, const void* A, const void* b, const void* G,
                                      const void* h, double solver_tol, int32_t flags, void* x, void* s, void* z,
                                      void* y, int32_t* converged, int32_t* iters, double* trace, void* workspace,
                                      size_t workspace_bytes, void* stream);

/* Elastic QP:  minimise 1/2 x'Qx + q'x + penalty 1't  s.t.  G x - t <= h, t >= 0 [, A x = b].
 * qpx_solve_qp_elastic replaces qpax.solve_qp_elastic (qpax/elastic_qp.py:184-228: the
 * initialisation :9-47 and the predictor-corrector loop of Algorithm 4 :99-174 with the
 * one-factorisation reduction of Algorithm 5 :50-96).  penalty is a device array of `batch`
 * values.  The equality rows (m_eq > 0: A, b in, y out) are an extension; the reference's elastic
 * functions take none (pass m_eq = 0 for the reference's API).
 * Outputs: x (B,n), t, s1, s2, z1, z2 (B,m_ineq), y (B,m_eq), converged, iters (B). */
QPX_API qpx_status qpx_solve_qp_elastic(int32_t dtype, int32_t mode, int64_t batch, int32_t n,
                                        int32_t m_eq, int32_t m_ineq, const void* Q, const void* q,
                                        const void* A, const void* b, const void* G, const void* h,
                                        const void* penalty, double solver_tol, void* x, void* t,
                                        void* s1, void* s2, void* z1, void* z2, void* y,
                                        int32_t* converged, int32_t* iters, void* workspace,
                                        size_t workspace_bytes, void* stream);

/* Forward of qpax.solve_qp_elastic_primal under differentiation -- replaces
 * solve_qp_elastic_primal_fwd (qpax/elastic_qp.py:433-456): solve, then relax_qp_elastic
 * (:319-359) from the solution.  flags: 0 or QPX_ELASTIC_CONSISTENT. */
QPX_API qpx_status qpx_solve_qp_elastic_relaxed(
    int32_t dtype, int32_t mode, int64_t batch, int32_t n, int32_t m_eq, int32_t m_ineq,
    const void* Q, const void* q, const void* A, const void* b, const void* G, const void* h,
    const void* penalty, double solver_tol, double target_kappa, int32_t flags, void* x, void* t,
    void* s1, void* s2, void* z1, void* z2, void* y, int32_t* converged, int32_t* iters, void* xr,
    void* tr, void* s1r, void* s2r, void* z1r, void* z2r, void* yr, int32_t* relax_converged,
    int32_t* relax_iters, void* workspace, size_t workspace_bytes, void* stream);

/* qpax.elastic_qp.initialize_elastic -- replaces qpax/elastic_qp.py:9-47 (solve_init_elastic_ls and
 * the CVXOPT shift on the stacked (z1, z2)), with Q symmetrised as solve_qp_elastic hands it over
 * (:186).  Outputs x (B,n), t, s1, s2, z1, z2 (B,m_ineq) [, y (B,m_eq) with equality rows]. */
QPX_API qpx_status qpx_initialize_elastic(int32_t dtype, int64_t batch, int32_t n, int32_t m_eq, int32_t m_ineq,
                                          const void* Q, const void* q, const void* A, const void* b, const void* G,
                                          const void* h, const void* penalty, void* x, void* t, void* s1, void* s2,
                                          void* z1, void* z2, void* y, void* workspace, size_t workspace_bytes,
                                          void* stream);

/* Backward of qpax.solve_qp_elastic_primal -- replaces solve_qp_elastic_primal_bwd /
 * diff_qp_elastic / optnet_derivatives_elastic (qpax/elastic_qp.py:366-417, :459-466):
 * cotangents dQ, dq, dG, dh (and dA, db with equality rows); none for the penalty.
 * flags: 0 or QPX_EXACT_DG. */
QPX_API qpx_status qpx_diff_qp_elastic(int32_t dtype, int32_t mode, int64_t batch, int32_t n,
                                       int32_t m_eq, int32_t m_ineq, const void* Q, const void* q,
                                       const void* A, const void* b, const void* G, const void* h,
                                       const void* penalty, const void* xr, const void* tr,
                                       const void* s1r, const void* s2r, const void* z1r,
                                       const void* z2r, const void* yr, const void* dl_dx,
                                       int32_t flags, void* dQ, void* dq, void* dA, void* db,
                                       void* dG, void* dh, void* workspace, size_t workspace_bytes,
                                       void* stream);

/* ---------------------------------- ragged batches ---------------------------------------------
 * Problems of different sizes in one call.  No counterpart in the reference, which batches one
 * size with jax.vmap over solve_qp / solve_qp_primal (qpax/pdip.py:181, qpax/diff_qp.py:50-79);
 * per problem the arithmetic is exactly that of qpx_solve_qp_relaxed / qpx_diff_qp in
 * QPX_FUSED mode.
 *
 * Operands are packed back to back in problem order: Q_0 (n_0 x n_0) then Q_1 ..., and likewise
 * q, A (m_eq_i x n_i), b, G (m_ineq_i x n_i), h.  qpx_plan_offsets gives each problem's element
 * offsets; outputs reuse them (x at q's offset, s and z at h's, y at b's, dQ at Q's, ...).
 * A plan is bound to the device current at creation and holds device tables for the size pattern;
 * calls through it allocate nothing and do not synchronise (capturable in a hipGraph).  Calls that
 * share one workspace must be stream-ordered.
 * A call runs its buckets on up to four lanes: the caller's stream and up to three side streams the
 * plan owns, forked from the caller's stream and joined back into it, so the call is still ordered
 * on the caller's stream as one unit (a captured call is a graph with one branch per lane).  The
 * environment variable QPX_RAGGED_LANES, read at plan creation, caps the lanes (1: every bucket on
 * the caller's stream).  Calls on one plan from several host threads are serialised while they
 * enqueue (a mutex per plan), so they may share a plan; each needs its own workspace. */
typedef struct qpx_plan qpx_plan;

/* Buckets the problems by kernel variant, orders each bucket longest-first and uploads the tables
 * (synchronous).  Every problem needs n >= 1 and m_ineq >= 1. */
QPX_API qpx_status qpx_plan_create(int32_t dtype, int64_t batch, const int32_t* n, const int32_t* m_eq,
                                   const int32_t* m_ineq, qpx_plan** plan);
QPX_API void qpx_plan_destroy(qpx_plan* plan);
/* info[0..5]: total elements of packed Q, q, A, b, G, h; info[6]: workspace bytes; info[7]: buckets. */
QPX_API qpx_status qpx_plan_info(const qpx_plan* plan, int64_t* info);
QPX_API size_t qpx_plan_workspace_size(const qpx_plan* plan);
/* offsets (batch x 6, host): element offsets of problem i in packed Q, q, A, b, G, h. */
QPX_API qpx_status qpx_plan_offsets(const qpx_plan* plan, int64_t* offsets);
/* out[0..9] of bucket i: kind (0 wavefront kernel, 1 generic workgroup kernel, 2 matrix-core
 * workgroup kernel), NMAX (kind 0) or tiles per side (kind 2), MIC, threads, wavefronts per
 * workgroup, persistent, problems, LDS bytes per wavefront (kind 0) or workgroup (kind 2),
 * workgroups, flops per iteration. */
QPX_API qpx_status qpx_plan_bucket(const qpx_plan* plan, int32_t i, int64_t* out);
/* qpx_solve_qp over a ragged batch (packed operands and outputs). */
QPX_API qpx_status qpx_plan_solve(const qpx_plan* plan, const void* Q, const void* q, const void* A,
                                  const void* b, const void* G, const void* h, double solver_tol,
                                  void* x, void* s, void* z, void* y, int32_t* converged,
                                  int32_t* iters, void* workspace, size_t workspace_bytes,
                                  void* stream);
/* qpx_solve_qp_relaxed over a ragged batch. */
QPX_API qpx_status qpx_plan_solve_relaxed(const qpx_plan* plan, const void* Q, const void* q,
                                          const void* A, const void* b, const void* G,
                                          const void* h, double solver_tol, double target_kappa,
                                          void* x, void* s, void* z, void* y, int32_t* converged,
                                          int32_t* iters, void* xr, void* sr, void* zr, void* yr,
                                          int32_t* relax_converged, int32_t* relax_iters,
                                          void* workspace, size_t workspace_bytes, void* stream);
/* qpx_diff_qp over a ragged batch. */
QPX_API qpx_status qpx_plan_diff(const qpx_plan* plan, const void* Q, const void* q, const void* A,
                                 const void* b, const void* G, const void* h, const void* xr,
                                 const void* sr, const void* zr, const void* yr, const void* dl_dx,
                                 int32_t flags, void* dQ, void* dq, void* dA, void* db, void* dG,
                                 void* dh, void* workspace, size_t workspace_bytes, void* stream);
/* Autotuner: times each bucket's candidate launch shapes (wavefronts per workgroup 1/2/4 or
 * threads 64/128/256, persistent queue or one workgroup per problem) with a solve + relax on
 * these operands, keeps the fastest in the plan, then runs the call once more with the choice.
 * Synchronises `stream`.  Same arguments as qpx_plan_solve_relaxed. */
QPX_API qpx_status qpx_plan_tune(qpx_plan* plan, const void* Q, const void* q, const void* A,
                                 const void* b, const void* G, const void* h, double solver_tol,
                                 double target_kappa, void* x, void* s, void* z, void* y,
                                 int32_t* converged, int32_t* iters, void* xr, void* sr, void* zr,
                                 void* yr, int32_t* relax_converged, int32_t* relax_iters,
                                 void* workspace, size_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* QPX_H_ */
