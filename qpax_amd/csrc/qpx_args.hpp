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
};

}  // namespace qpx
