// Argument block shared by every kernel of libqpx (passed by value at launch).
#pragma once
#include <stdint.h>

namespace qpx {

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
};

}  // namespace qpx
