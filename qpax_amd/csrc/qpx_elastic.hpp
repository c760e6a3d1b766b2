// Elastic (soft-constraint) QP -- qpax/elastic_qp.py -- on the generic block-per-QP path.
//
//   minimise 1/2 x'Qx + q'x + penalty 1't   s.t.  G x - t <= h,  t >= 0   [, A x = b]
//
// Variables x (n), t, s1, s2, z1, z2 (m_ineq each) [, y (m_eq)].  One n x n factorisation per
// iteration, H = Q + G^T diag(1/a3) G with a3 = s1/z1 + s2/z2 (Algorithm 5,
// qpax/elastic_qp.py:50-96), reused by the corrector.  The equality rows are an extension (the
// reference's elastic functions take no A, b): with them the single H solve of Algorithm 5
// becomes the reduced KKT solve of [[H, A^T], [A, 0]], exactly as the standard form does it.
#pragma once
#include "qpx_args.hpp"
#include "qpx_core.hpp"

namespace qpx {

// Elastic extras beyond the standard View (which provides Q, G, A, q, b, h, x, y, D, H, W, F,
// r1 (stationarity), r4 (equality rows), dx, dy, u, t (m_eq scratch)).
struct EView {
  View v;
  double *tt, *s1, *s2, *z1, *z2;           // iterate (tt: the elastic slack t)
  double *r2, *r3, *r4e, *r5, *r6;         // residual blocks 2..6 (r1 and the equality block in v)
  double *dt, *ds1, *ds2, *dz1, *dz2;      // direction
  double *ds1a, *ds2a, *dz1a, *dz2a, *p1;  // affine direction, Algorithm 5's p1
  double* pa3;                             // p1 / a3
  double pen;
};

__host__ __device__ inline int64_t elastic_words(int n, int me, int mi) {
  return view_words(n, me, mi) + round_up(21 * round_up(mi > 0 ? mi : 1, 2), 64);
}

__host__ __device__ inline EView carve_elastic(double* base, int n, int me, int mi) {
  EView e;
  e.v = carve(base, n, me, mi);
  double* p = base + view_words(n, me, mi);
  const int64_t st = round_up(mi > 0 ? mi : 1, 2);
  // member by member (taking the members' addresses would keep the whole view in scratch memory)
  auto take = [&]() { double* r = p; p += st; return r; };
  e.tt = take(); e.s1 = take(); e.s2 = take(); e.z1 = take(); e.z2 = take();
  e.r2 = take(); e.r3 = take(); e.r4e = take(); e.r5 = take(); e.r6 = take();
  e.dt = take(); e.ds1 = take(); e.ds2 = take(); e.dz1 = take(); e.dz2 = take();
  e.ds1a = take(); e.ds2a = take(); e.dz1a = take(); e.dz2a = take(); e.p1 = take();
  e.pa3 = take();
  e.pen = 0.0;
  return e;
}

// Load (qpax/elastic_qp.py:186 symmetrisation; no inf-bound handling in the elastic reference).
template <int NT, typename TS>
__device__ void e_load(EView& e, const TS* Q, const TS* q, const TS* A, const TS* b, const TS* G,
                       const TS* h, const TS* pen) {
  const View& v = e.v;
  const int n = v.n, me = v.me, mi = v.mi, tid = threadIdx.x;
  for (int i = tid; i < n * n; i += NT) {
    const int r = i / n, c = i - r * n;
    v.Q[i] = 0.5 * ((double)Q[i] + (double)Q[c * n + r]);
  }
  for (int i = tid; i < mi * n; i += NT) v.G[i] = (double)G[i];
  for (int i = tid; i < me * n; i += NT) v.A[i] = (double)A[i];
  for (int i = tid; i < n; i += NT) v.q[i] = (double)q[i];
  for (int c = tid; c < me; c += NT) v.b[c] = (double)b[c];
  for (int k = tid; k < mi; k += NT) v.h[k] = (double)h[k];
  if (tid < kCtlWords) v.ctl[tid] = 0;
  e.pen = (double)pen[0];
  __syncthreads();
}

// Initialisation (qpax/elastic_qp.py:9-47): H0 = Q + G^T G / 2,
// x = H0^{-1}(-q - G^T (pen - h) / 2), z2 = (G x + pen - h) / 2, z1 = pen - z2, t = -z1, then the
// CVXOPT shift on the stacked (z1, z2).
// PF: the sweeps with L loaded ahead (block_fsub_tiled), for the one-workgroup-per-CU fused kernels
template <int NT, bool PF = false>
__device__ void e_init(EView& e, double* red) {
  const View& v = e.v;
  const int n = v.n, me = v.me, mi = v.mi, tid = threadIdx.x;
  for (int k = tid; k < mi; k += NT) v.D[k] = 0.5;
  stage_factor<NT>(v, kDgiven);
  for (int i = tid; i < n; i += NT) {
    double acc = 0.0;
    for (int k = 0; k < mi; ++k) acc += v.G[k * n + i] * (e.pen - v.h[k]);
    v.u[i] = -v.q[i] - 0.5 * acc;
  }
  __syncthreads();
  reduced_solve<NT, PF>(v, v.b, v.x);
  for (int c = tid; c < me; c += NT) v.y[c] = v.dy[c];
  double zmax = -INFINITY, zmin = INFINITY;
  for (int k = tid; k < mi; k += NT) {
    double g = 0.0;
    for (int j = 0; j < n; ++j) g += v.G[k * n + j] * v.x[j];
    const double z2 = 0.5 * (g + e.pen - v.h[k]);
    const double z1 = e.pen - z2;
    e.tt[k] = -z1;
    e.z1[k] = z1;
    e.z2[k] = z2;
    zmax = nmax(zmax, nmax(z1, z2));
    zmin = nmin(zmin, nmin(z1, z2));
  }
  zmax = block_reduce<NT>(zmax, red, OpMax());
  zmin = block_reduce<NT>(zmin, red, OpMin());
  const double ap = zmax, ad = -zmin;
  for (int k = tid; k < mi; k += NT) {
    const double z1 = e.z1[k], z2 = e.z2[k];
    e.s1[k] = (ap < 0) ? -z1 : -z1 + (1 + ap);
    e.s2[k] = (ap < 0) ? -z2 : -z2 + (1 + ap);
    e.z1[k] = (ad >= 0) ? z1 + (1 + ad) : z1;
    e.z2[k] = (ad >= 0) ? z2 + (1 + ad) : z2;
  }
  __syncthreads();
}

// Residuals (qpax/elastic_qp.py:105-110; relaxed: :256-261).  The reference's relaxed residuals
// flip the sign of t and of r2 (sign_t = +1, sign_r2 = -1); the forward uses sign_t = -1,
// sign_r2 = +1.  Returns the entry convergence flag.
template <int NT, typename GT, bool PF = false>
__device__ int e_residuals(EView& e, double kappa, double sign_t, double sign_r2, double tol, double* red,
                           const GT* G) {
  const View& v = e.v;
  const int n = v.n, me = v.me, mi = v.mi, tid = threadIdx.x;
  const bool wide = n >= kMfmaMinN;
  // PF (the fused kernels): G^T z2 from the same pass over G as G x (rows_cols16), kept in LDS past
  // the partial sums of both passes; the per-stage kernels keep two passes and fewer registers
  constexpr int CPL = colsum_slots<NT>();
  QPX_DYN_LDS(qpx_smem);
  double* part = reinterpret_cast<double*>(qpx_smem);
  const int gofs = 4 * (NT / 64) * 16 * ((n + 15) / 16), cofs = (NT / 64) * 64 * CPL;
  double* gtz = part + (gofs > cofs ? gofs : cofs);
  const bool fuse = PF && wide && n <= kMfmaMaxN;
  if (wide) {  // G x into r6 and A x into r4 first; Q x by columns (Q is symmetric), coalesced
    if constexpr (PF) {
      if (n <= kMfmaMaxN) rows_cols16<NT>(G, mi, n, v.x, e.r6, e.z2, gtz, part);
      else rows_dot16<NT>(G, mi, n, v.x, e.r6);
    } else {
      rows_dot16<NT>(G, mi, n, v.x, e.r6);
    }
    rows_dot16<NT>(v.A, me, n, v.x, v.r4);
    __syncthreads();
  }
  double m = 0.0;
  if (wide && n <= kMfmaMaxN) {  // r1 = q + Q x + G^T z2 + A^T y as column sums (Q is symmetric)
    double acc[CPL] = {};
    colsum_acc<NT, CPL>(v.Q, n, n, v.x, acc);
    if (!fuse) colsum_acc<NT, CPL>(G, mi, n, e.z2, acc);
    colsum_acc<NT, CPL>(v.A, me, n, v.y, acc);
    colsum_finish<NT, CPL>(acc, n, v.q, 1.0, 1.0, v.r1, part);
    for (int i = tid; i < n; i += NT) {
      const double r1 = fuse ? v.r1[i] + gtz[i] : v.r1[i];
      v.r1[i] = r1;
      m = nmax(m, fabs(r1));
    }
  } else {
    for (int i = tid; i < n; i += NT) {
      double acc = v.q[i];
      if (wide)
        for (int j = 0; j < n; ++j) acc += v.Q[j * n + i] * v.x[j];
      else
        for (int j = 0; j < n; ++j) acc += v.Q[i * n + j] * v.x[j];
      for (int k = 0; k < mi; ++k) acc += (double)G[k * n + i] * e.z2[k];
      for (int c = 0; c < me; ++c) acc += v.A[c * n + i] * v.y[c];
      v.r1[i] = acc;
      m = nmax(m, fabs(acc));
    }
  }
  for (int k = tid; k < mi; k += NT) {
    double g = 0.0;
    if (wide) g = e.r6[k];
    else
      for (int j = 0; j < n; ++j) g += (double)G[k * n + j] * v.x[j];
    const double r2 = sign_r2 * (-e.z1[k] - e.z2[k] + e.pen);
    const double r3 = e.s1[k] * e.z1[k] - kappa;
    const double r4 = e.s2[k] * e.z2[k] - kappa;
    const double r5 = sign_t * e.tt[k] + e.s1[k];
    const double r6 = g + sign_t * e.tt[k] + e.s2[k] - v.h[k];
    e.r2[k] = r2; e.r3[k] = r3; e.r4e[k] = r4; e.r5[k] = r5; e.r6[k] = r6;
    m = nmax(m, nmax(nmax(fabs(r2), fabs(r3)), nmax(nmax(fabs(r4), fabs(r5)), fabs(r6))));
  }
  for (int c = tid; c < me; c += NT) {
    double acc = -v.b[c];
    if (wide) acc += v.r4[c];
    else
      for (int j = 0; j < n; ++j) acc += v.A[c * n + j] * v.x[j];
    v.r4[c] = acc;
    m = nmax(m, fabs(acc));
  }
  m = block_reduce<NT>(m, red, OpMax());
  __syncthreads();
  return (m < tol) ? 1 : 0;
}

// Factorisation for Algorithm 5 at (s1, z1, s2, z2): D = 1/a3.
template <int NT, typename GT>
__device__ void e_factor(EView& e, const GT* G) {
  const View& v = e.v;
  for (int k = threadIdx.x; k < v.mi; k += NT) v.D[k] = 1.0 / (e.s1[k] / e.z1[k] + e.s2[k] / e.z2[k]);
  stage_factor_g<NT, GT>(v, kDgiven, G);
}

// Algorithm 5 with right-hand side sg * (R1, R2, R3, R4, R5, R6, R7) (R7: the equality rows) using
// the current factorisation.  Writes dx, dy (in v) and dt, ds1, ds2, dz1, dz2 (given arrays).
template <int NT, typename GT, bool PF = false>
__device__ void e_solve(EView& e, const double* R1, const double* R2, const double* R3, const double* R4,
                        const double* R5, const double* R6, const double* R7, double sg, double* dt,
                        double* ds1, double* ds2, double* dz1, double* dz2, const GT* G) {
  const View& v = e.v;
  const int n = v.n, me = v.me, mi = v.mi, tid = threadIdx.x;
  for (int k = tid; k < mi; k += NT) {
    const double a1 = e.s1[k] / e.z1[k], a2 = e.s2[k] / e.z2[k];
    const double w1 = sg * R3[k] / e.z1[k], w2 = sg * R4[k] / e.z2[k];
    const double p1 = sg * R5[k] - sg * R6[k] + w2 - w1 - a1 * (sg * R2[k]);
    e.p1[k] = p1;
    e.pa3[k] = p1 / (a1 + a2);
  }
  __syncthreads();
  if (n >= kMfmaMinN && n <= kMfmaMaxN) {  // u = sg R1 - G^T (p1 / a3) as column sums
    constexpr int CPL = colsum_slots<NT>();
    QPX_DYN_LDS(qpx_smem);
    double acc[CPL] = {};
    colsum_acc<NT, CPL>(G, mi, n, e.pa3, acc);
    colsum_finish<NT, CPL>(acc, n, R1, sg, -1.0, v.u, reinterpret_cast<double*>(qpx_smem));
  } else {
    for (int i = tid; i < n; i += NT) {
      double acc = 0.0;
      for (int k = 0; k < mi; ++k) acc += (double)G[k * n + i] * e.pa3[k];
      v.u[i] = sg * R1[i] - acc;
    }
  }
  for (int c = tid; c < me; c += NT) v.dy[c] = sg * R7[c];
  __syncthreads();
  reduced_solve<NT, PF>(v, v.dy, v.dx);
  const bool wide = n >= kMfmaMinN;
  if (wide) {  // G dx into dz2 first
    rows_dot16<NT>(G, mi, n, v.dx, dz2);
    __syncthreads();
  }
  for (int k = tid; k < mi; k += NT) {
    double g = 0.0;
    if (wide) g = dz2[k];
    else
      for (int j = 0; j < n; ++j) g += (double)G[k * n + j] * v.dx[j];
    const double a3 = e.s1[k] / e.z1[k] + e.s2[k] / e.z2[k];
    const double d2 = (e.p1[k] + g) / a3;
    const double d1 = -(sg * R2[k]) - d2;
    const double s1v = (sg * R3[k] - e.s1[k] * d1) / e.z1[k];
    const double s2v = (sg * R4[k] - e.s2[k] * d2) / e.z2[k];
    dz2[k] = d2;
    dz1[k] = d1;
    ds1[k] = s1v;
    ds2[k] = s2v;
    dt[k] = s1v - sg * R5[k];
  }
  __syncthreads();
}

// Fraction-to-boundary over the stacked (v1; v2) (qpax/pdip.py:66-73).
template <int NT>
__device__ double e_ort(const double* x1, const double* d1, const double* x2, const double* d2, int mi, double* red) {
  double a = 1.0;
  for (int k = threadIdx.x; k < mi; k += NT) {
    if (d1[k] < 0) a = nmin(a, -x1[k] / d1[k]);
    if (d2[k] < 0) a = nmin(a, -x2[k] / d2[k]);
  }
  return block_reduce<NT>(a, red, OpMin());
}

// Centering on the stacked (s1; s2), (z1; z2) and the corrector right-hand side
// (qpax/elastic_qp.py:119-127).
template <int NT>
__device__ void e_center(EView& e, double* red) {
  const int mi = e.v.mi, tid = threadIdx.x;
  double sz = 0.0;
  for (int k = tid; k < mi; k += NT) sz += e.s1[k] * e.z1[k] + e.s2[k] * e.z2[k];
  sz = block_reduce<NT>(sz, red, OpSum());
  const double mu = sz / (2 * mi);
  const double a = nmin(e_ort<NT>(e.s1, e.ds1a, e.s2, e.ds2a, mi, red), e_ort<NT>(e.z1, e.dz1a, e.z2, e.dz2a, mi, red));
  double g = 0.0;
  for (int k = tid; k < mi; k += NT)
    g += (e.s1[k] + a * e.ds1a[k]) * (e.z1[k] + a * e.dz1a[k]) + (e.s2[k] + a * e.ds2a[k]) * (e.z2[k] + a * e.dz2a[k]);
  g = block_reduce<NT>(g, red, OpSum());
  const double ratio = g / sz;
  const double sm = ratio * ratio * ratio * mu;
  for (int k = tid; k < mi; k += NT) {
    e.r3[k] = e.r3[k] - (sm - e.ds1a[k] * e.dz1a[k]);
    e.r4e[k] = e.r4e[k] - (sm - e.ds2a[k] * e.dz2a[k]);
  }
  __syncthreads();
}

// Step with alpha = 0.99 min(1, 0.99 ort(s), 0.99 ort(z)) (qpax/elastic_qp.py:135-142, :276-283).
template <int NT>
__device__ void e_step(EView& e, double* red) {
  const View& v = e.v;
  const int n = v.n, me = v.me, mi = v.mi, tid = threadIdx.x;
  const double as = e_ort<NT>(e.s1, e.ds1, e.s2, e.ds2, mi, red);
  const double az = e_ort<NT>(e.z1, e.dz1, e.z2, e.dz2, mi, red);
  const double alpha = 0.99 * nmin(nmin(1.0, 0.99 * as), 0.99 * az);
  for (int i = tid; i < n; i += NT) v.x[i] += alpha * v.dx[i];
  for (int c = tid; c < me; c += NT) v.y[c] += alpha * v.dy[c];
  for (int k = tid; k < mi; k += NT) {
    e.tt[k] += alpha * e.dt[k];
    e.s1[k] += alpha * e.ds1[k];
    e.s2[k] += alpha * e.ds2[k];
    e.z1[k] += alpha * e.dz1[k];
    e.z2[k] += alpha * e.dz2[k];
  }
  __syncthreads();
}

// One predictor-corrector iteration (qpax/elastic_qp.py:99-174); returns the entry flag.
template <int NT, typename GT>
__device__ int e_pc_iteration(EView& e, double tol, double* red, const GT* G) {
  const int conv = e_residuals<NT, GT, true>(e, 0.0, -1.0, 1.0, tol, red, G);
  e_factor<NT>(e, G);
  e_solve<NT, GT, true>(e, e.v.r1, e.r2, e.r3, e.r4e, e.r5, e.r6, e.v.r4, -1.0, e.dt, e.ds1a, e.ds2a, e.dz1a, e.dz2a, G);
  e_center<NT>(e, red);
  e_solve<NT, GT, true>(e, e.v.r1, e.r2, e.r3, e.r4e, e.r5, e.r6, e.v.r4, -1.0, e.dt, e.ds1, e.ds2, e.dz1, e.dz2, G);
  e_step<NT>(e, red);
  return conv;
}

// One relaxed Newton iteration (qpax/elastic_qp.py:235-316); consistent != 0 uses the forward's
// signs of t and r2.
template <int NT, typename GT>
__device__ int e_relax_iteration(EView& e, double tol, double kappa, int consistent, double* red, const GT* G) {
  const int conv = consistent ? e_residuals<NT, GT, true>(e, kappa, -1.0, 1.0, tol, red, G)
                              : e_residuals<NT, GT, true>(e, kappa, 1.0, -1.0, tol, red, G);
  e_factor<NT>(e, G);
  e_solve<NT, GT, true>(e, e.v.r1, e.r2, e.r3, e.r4e, e.r5, e.r6, e.v.r4, -1.0, e.dt, e.ds1, e.ds2, e.dz1, e.dz2, G);
  e_step<NT>(e, red);
  return conv;
}

template <typename TS>
__device__ __forceinline__ EView eview_of(const Args<TS>& a, int64_t p) {
  EView e = carve_elastic(a.ws + p * a.stride, a.n, a.me, a.mi);
  return e;
}

template <int NT, typename TS>
__device__ void e_write_state(EView& e, int64_t p, TS* x, TS* t, TS* s1, TS* s2, TS* z1, TS* z2, TS* y) {
  const View& v = e.v;
  const int n = v.n, me = v.me, mi = v.mi, tid = threadIdx.x;
  for (int i = tid; i < n; i += NT) x[p * n + i] = (TS)v.x[i];
  for (int c = tid; c < me; c += NT) y[p * me + c] = (TS)v.y[c];
  for (int k = tid; k < mi; k += NT) {
    t[p * mi + k] = (TS)e.tt[k];
    s1[p * mi + k] = (TS)e.s1[k];
    s2[p * mi + k] = (TS)e.s2[k];
    z1[p * mi + k] = (TS)e.z1[k];
    z2[p * mi + k] = (TS)e.z2[k];
  }
}

// ---------------------------------- kernels -------------------------------------------------
#define QPX_E_LOAD(NT_, TS_, e, a, p)                                                                   \
  e_load<NT_, TS_>(e, a.Q + (p) * a.n * a.n, a.q + (p) * a.n, a.A + (p) * a.me * a.n, a.b + (p) * a.me, \
                   a.G + (p) * a.mi * a.n, a.h + (p) * a.mi, a.pen + (p))

template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_e_setup(Args<TS> a) {
  __shared__ double red[NT / 64];
  const int64_t p = blockIdx.x;
  EView e = eview_of(a, p);
  QPX_E_LOAD(NT, TS, e, a, p);
  e_init<NT>(e, red);
}

// initialize_elastic (qpax/elastic_qp.py:9-47) alone, written to x, t, s1, s2, z1, z2 (and y).
template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_e_initialize(Args<TS> a) {
  __shared__ double red[NT / 64];
  const int64_t p = blockIdx.x;
  EView e = eview_of(a, p);
  QPX_E_LOAD(NT, TS, e, a, p);
  e_init<NT>(e, red);
  e_write_state<NT, TS>(e, p, a.x, a.t, a.s, a.s2, a.z, a.z2, a.y);
}

// Plain mode: one kernel per iteration-stage group (residuals, factor, solves, step).
template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_e_resid(Args<TS> a, int relax) {
  __shared__ double red[NT / 64];
  if (blockIdx.x == 0 && threadIdx.x == 0) *a.counter = 0u;
  EView e = eview_of(a, blockIdx.x);
  e.pen = (double)a.pen[blockIdx.x];
  const View& v = e.v;
  const int c = v.ctl[relax ? kRelaxConv : kConv], it = v.ctl[relax ? kRelaxIter : kIter];
  const bool act = it < kMaxIters && c == 0;
  __syncthreads();
  if (threadIdx.x == 0) v.ctl[kActive] = act ? 1 : 0;
  if (!act) return;
  const bool fwd_signs = !relax || (a.flags & 2);
  const int conv = fwd_signs ? e_residuals<NT>(e, relax ? a.kappa : 0.0, -1.0, 1.0, a.tol, red, e.v.G)
                             : e_residuals<NT>(e, a.kappa, 1.0, -1.0, a.tol, red, e.v.G);
  if (threadIdx.x == 0) v.ctl[5] = conv;
}

template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_e_factor(Args<TS> a) {
  EView e = eview_of(a, blockIdx.x);
  if (!e.v.ctl[kActive]) return;
  e_factor<NT>(e, e.v.G);
}

template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_e_solve(Args<TS> a, int affine) {
  EView e = eview_of(a, blockIdx.x);
  if (!e.v.ctl[kActive]) return;
  if (affine) e_solve<NT>(e, e.v.r1, e.r2, e.r3, e.r4e, e.r5, e.r6, e.v.r4, -1.0, e.dt, e.ds1a, e.ds2a, e.dz1a, e.dz2a, e.v.G);
  else e_solve<NT>(e, e.v.r1, e.r2, e.r3, e.r4e, e.r5, e.r6, e.v.r4, -1.0, e.dt, e.ds1, e.ds2, e.dz1, e.dz2, e.v.G);
}

template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_e_center(Args<TS> a) {
  __shared__ double red[NT / 64];
  EView e = eview_of(a, blockIdx.x);
  if (!e.v.ctl[kActive]) return;
  e_center<NT>(e, red);
}

template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_e_step(Args<TS> a, int relax) {
  __shared__ double red[NT / 64];
  EView e = eview_of(a, blockIdx.x);
  const View& v = e.v;
  if (!v.ctl[kActive]) return;
  e_step<NT>(e, red);
  if (threadIdx.x == 0) {
    v.ctl[relax ? kRelaxConv : kConv] = v.ctl[5];
    v.ctl[relax ? kRelaxIter : kIter] += 1;
    const int c = v.ctl[relax ? kRelaxConv : kConv], it = v.ctl[relax ? kRelaxIter : kIter];
    if (it < kMaxIters && c == 0) atomicAdd(a.counter, 1u);
  }
}

template <int NT, typename TS>
__device__ void e_finish_solve(const Args<TS>& a, EView& e, int64_t p) {
  e_write_state<NT, TS>(e, p, a.x, a.t, a.s, a.s2, a.z, a.z2, a.y);
  __syncthreads();
  if (threadIdx.x == 0) {
    a.conv[p] = e.v.ctl[kConv];
    a.iters[p] = e.v.ctl[kIter];
    e.v.ctl[kRelaxConv] = 0;
    e.v.ctl[kRelaxIter] = 0;
  }
  __syncthreads();
}

template <int NT, typename TS>
__device__ void e_finish_relax(const Args<TS>& a, EView& e, int64_t p) {
  e_write_state<NT, TS>(e, p, a.xr, a.tr, a.sr, a.s2r, a.zr, a.z2r, a.yr);
  if (threadIdx.x == 0) { a.rconv[p] = e.v.ctl[kRelaxConv]; a.riters[p] = e.v.ctl[kRelaxIter]; }
}

template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_e_finish_solve(Args<TS> a) {
  EView e = eview_of(a, blockIdx.x);
  e_finish_solve<NT, TS>(a, e, blockIdx.x);
}

template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_e_finish_relax(Args<TS> a) {
  EView e = eview_of(a, blockIdx.x);
  e_finish_relax<NT, TS>(a, e, blockIdx.x);
}

// Fused mode: every block iterates its own problem to completion.
template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_e_fused_solve(Args<TS> a, int with_relax) {
  __shared__ double red[NT / 64];
  const int64_t p = blockIdx.x;
  EView e = eview_of(a, p);
  QPX_E_LOAD(NT, TS, e, a, p);
  e_init<NT, true>(e, red);
  // G straight from the caller's array (the workspace copy holds the same values widened to
  // float64): half the bytes of the per-iteration G passes for float32 inputs
  const TS* Gin = a.G + p * a.mi * a.n;
  int conv = 0, it = 0;
  while (it < kMaxIters && conv == 0) {
    conv = e_pc_iteration<NT>(e, a.tol, red, Gin);
    ++it;
  }
  if (threadIdx.x == 0) { e.v.ctl[kConv] = conv; e.v.ctl[kIter] = it; }
  __syncthreads();
  e_finish_solve<NT, TS>(a, e, p);
  if (!with_relax) return;
  int rconv = 0, rit = 0;
  while (rit < kMaxIters && rconv == 0) {
    rconv = e_relax_iteration<NT>(e, a.tol, a.kappa, a.flags & 2, red, Gin);
    ++rit;
  }
  if (threadIdx.x == 0) { e.v.ctl[kRelaxConv] = rconv; e.v.ctl[kRelaxIter] = rit; }
  __syncthreads();
  e_finish_relax<NT, TS>(a, e, p);
}

// Backward (qpax/elastic_qp.py:366-417): factor at the relaxed point, Algorithm 5 with right-hand
// side (-dl/dx, 0, ..., 0), cotangents of (Q, q, G, h) [and (A, b)].
template <int NT, typename TS>
__device__ void e_bwd_load(const Args<TS>& a, EView& e, int64_t p) {
  QPX_E_LOAD(NT, TS, e, a, p);
  const View& v = e.v;
  const int n = v.n, me = v.me, mi = v.mi, tid = threadIdx.x;
  for (int i = tid; i < n; i += NT) { v.x[i] = (double)a.xr_in[p * n + i]; v.r1[i] = (double)a.g[p * n + i]; }
  for (int c = tid; c < me; c += NT) { v.y[c] = (double)a.yr_in[p * me + c]; v.r4[c] = 0.0; }
  for (int k = tid; k < mi; k += NT) {
    e.tt[k] = (double)a.tr_in[p * mi + k];
    e.s1[k] = (double)a.sr_in[p * mi + k];
    e.z1[k] = (double)a.zr_in[p * mi + k];
    e.s2[k] = (double)a.s2r_in[p * mi + k];
    e.z2[k] = (double)a.z2r_in[p * mi + k];
    e.r2[k] = 0.0; e.r3[k] = 0.0; e.r4e[k] = 0.0; e.r5[k] = 0.0; e.r6[k] = 0.0;
  }
  __syncthreads();
}

template <int NT, typename TS, bool PF = false>
__device__ void e_bwd_solve_write(const Args<TS>& a, EView& e, int64_t p) {
  const View& v = e.v;
  const int n = v.n, me = v.me, mi = v.mi, tid = threadIdx.x;
  e_solve<NT, double, PF>(e, v.r1, e.r2, e.r3, e.r4e, e.r5, e.r6, v.r4, -1.0, e.dt, e.ds1, e.ds2, e.dz1, e.dz2, v.G);
  const bool exact = a.flags & 1;
  // x, dx, y, dy and each row's lam, dlam (dlt with `exact`) in LDS, read as broadcasts; a thread
  // writes one column of dQ, dG and dA (coalesced along the rows, no division per element).  Where
  // the kernel's dynamic LDS is too small for them (small problems), one element per thread.
  if (8 * (2 * n + 2 * me + 2 * mi) > generic_lds_bytes(n, me, NT)) {
    for (int i = tid; i < n * n; i += NT) {
      const int r = i / n, c = i - r * n;
      a.dQ[p * n * n + i] = (TS)(0.5 * (v.dx[r] * v.x[c] + v.x[r] * v.dx[c]));
    }
    for (int i = tid; i < mi * n; i += NT) {
      const int k = i / n, j = i - k * n;
      const double lam = e.z2[k], dlt = e.dz2[k], dlam = dlt / lam;
      const double val = exact ? dlt * v.x[j] + lam * v.dx[j] : lam * (dlam * v.x[j] + lam * v.dx[j]);
      a.dG[(int64_t)p * mi * n + i] = (TS)val;
    }
    for (int i = tid; i < me * n; i += NT) {
      const int c = i / n, j = i - c * n;
      a.dA[(int64_t)p * me * n + i] = (TS)(v.dy[c] * v.x[j] + v.y[c] * v.dx[j]);
    }
  } else {
    QPX_DYN_LDS(qpx_smem);
    double* sx = reinterpret_cast<double*>(qpx_smem);
    double* sdx = sx + n;
    double* sy = sdx + n;
    double* sdy = sy + me;
    double* slam = sdy + me;
    double* sdl = slam + mi;
    __syncthreads();
    for (int i = tid; i < n; i += NT) {
      sx[i] = v.x[i];
      sdx[i] = v.dx[i];
    }
    for (int c = tid; c < me; c += NT) {
      sy[c] = v.y[c];
      sdy[c] = v.dy[c];
    }
    for (int k = tid; k < mi; k += NT) {
      const double lam = e.z2[k], dlt = e.dz2[k];
      slam[k] = lam;
      sdl[k] = exact ? dlt : dlt / lam;
    }
    __syncthreads();
    for (int c = tid; c < n; c += NT) {
      const double xc = sx[c], dxc = sdx[c];
      for (int r = 0; r < n; ++r) a.dQ[(int64_t)p * n * n + r * n + c] = (TS)(0.5 * (sdx[r] * xc + sx[r] * dxc));
      for (int k = 0; k < mi; ++k) {
        const double lam = slam[k], d = sdl[k];
        const double val = exact ? d * xc + lam * dxc : lam * (d * xc + lam * dxc);
        a.dG[(int64_t)p * mi * n + k * n + c] = (TS)val;
      }
      for (int cc = 0; cc < me; ++cc) a.dA[(int64_t)p * me * n + cc * n + c] = (TS)(sdy[cc] * xc + sy[cc] * dxc);
    }
  }
  for (int i = tid; i < n; i += NT) a.dq[p * n + i] = (TS)v.dx[i];
  for (int k = tid; k < mi; k += NT) a.dh[p * mi + k] = (TS)(-e.z2[k] * (e.dz2[k] / e.z2[k]));
  for (int c = tid; c < me; c += NT) a.db[p * me + c] = (TS)(-v.dy[c]);
}

template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_e_bwd_load(Args<TS> a) {
  EView e = eview_of(a, blockIdx.x);
  e_bwd_load<NT, TS>(a, e, blockIdx.x);
}

template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_e_bwd_factor(Args<TS> a) {
  EView e = eview_of(a, blockIdx.x);
  e_factor<NT>(e, e.v.G);
}

template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_e_bwd_solve(Args<TS> a) {
  EView e = eview_of(a, blockIdx.x);
  e_bwd_solve_write<NT, TS>(a, e, blockIdx.x);
}

template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_e_fused_bwd(Args<TS> a) {
  const int64_t p = blockIdx.x;
  EView e = eview_of(a, p);
  e_bwd_load<NT, TS>(a, e, p);
  e_factor<NT>(e, e.v.G);
  e_bwd_solve_write<NT, TS, true>(a, e, p);
}

}  // namespace qpx
