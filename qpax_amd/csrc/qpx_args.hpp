// Argument block shared by every kernel of libqpx (passed by value at launch).
#pragma once
#include <stdint.h>

namespace qpx {

// One problem of a ragged batch: its sizes and the element offsets of its slices in the packed
// operand buffers (outputs reuse them: x and dq at off_q, s, z and dh at off_h, y and db at off_b,
// dQ at off_Q, dA at off_A, dG at off_G) and the float64-word offset of its workspace slice.
struct RaggedEntry {
  int32_t n, me, mi, reserved;
  int64_t off_Q, off_q, off_A, off_b, off_G, off_h, off_ws;
};

template <typename TS>
struct Args {
  int64_t B;
  int n, me, mi;
  const TS *Q, *q, *A, *b, *G, *h;
  TS *x, *s, *z, *y;
  int *conv, *iters;
  TS *xr, *sr, *zr, *yr;
  int *rconv, *riters;
  const TS *xr_in, *sr_in, *zr_in, *yr_in, *g;
  TS *dQ, *dq, *dA, *db, *dG, *dh;
  double* ws;
  int64_t stride;  // float64 words per problem
  unsigned* counter;
  double tol, kappa;
  int flags;       // QPX_EXACT_DG, QPX_ELASTIC_CONSISTENT (include/qpx.h)
  // elastic QP (s, z, sr, zr and their inputs above hold s1, z1)
  const TS* pen;
  TS *t, *s2, *z2, *tr, *s2r, *z2r;
  const TS *tr_in, *s2r_in, *z2r_in;
  // ragged batches and persistent scheduling
  const RaggedEntry* rg;  // per-problem sizes and offsets (null: uniform batch of size n, me, mi)
  const int* order;       // problems of this launch, in scheduling order (null: 0..B-1)
  int order_count;
  unsigned* qhead;        // work-queue head for persistent launches (null: one problem per slot)
  int lds_per_wave;       // bytes of LDS per wavefront (wavefront-per-QP kernels)
  int64_t split_fused;    // split forward launch: problems whose relaxation runs in the solve's item
  int split_spin;         // split forward launch: polls of a relaxation's wait for its solve
  double* trace;          // per-problem iteration trace (kMaxIters rows of kTraceCols), or null
};

constexpr int kTraceCols = 5;  // |r1|, |r2|, |r3|, |r4| (infinity norms at entry), step length

// Sizes and slice offsets of problem p.
struct PIdx {
  int n, me, mi;
  int64_t oQ, oq, oA, ob, oG, oh, ows;
};

template <typename TS>
__host__ __device__ inline PIdx pidx(const Args<TS>& a, int64_t p) {
  if (a.rg) {
    const RaggedEntry& e = a.rg[p];
    return PIdx{e.n, e.me, e.mi, e.off_Q, e.off_q, e.off_A, e.off_b, e.off_G, e.off_h, e.off_ws};
  }
  const int64_t n = a.n, me = a.me, mi = a.mi;
  return PIdx{a.n, a.me, a.mi, p * n * n, p * n, p * me * n, p * me, p * mi * n, p * mi, p * a.stride};
}

}  // namespace qpx
