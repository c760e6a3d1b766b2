// Kernels of the generic (any size) solver path and the C ABI of libqpx.
//
// One thread block owns one QP.  The plain mode launches one kernel per IPM stage per iteration
// over the whole padded batch and reads the number of still-active problems back to the host
// after every iteration; the fused mode runs each phase (solve + relax, backward) as a single
// launch in which every block iterates its own problem to completion.  The fused mode of a size
// that neither the wavefront nor the matrix-core workgroup kernels take calls the same device
// stages in the same order as the plain mode, bit for bit; for the other sizes the fused kernels
// run another elimination and agree with the plain mode to rounding.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>

#include "../../include/qpx.h"
#include "qpx_args.hpp"
#include "qpx_host.hpp"
#include "qpx_core.hpp"
#include "qpx_elastic.hpp"

#include "qpx_wave_launch.hpp"
#include "qpx_group.hpp"
namespace qpx {

template <typename TS>
__device__ __forceinline__ View view_of(const Args<TS>& a, const PIdx& pi) {
  return carve(a.ws + pi.ows, pi.n, pi.me, pi.mi);
}

template <typename TS>
__device__ __forceinline__ View view_of(const Args<TS>& a, int64_t p) {
  return view_of(a, pidx(a, p));
}

// Next problem for the calling workgroup: entry blockIdx.x of the schedule once, or, with a work
// queue, the entry whose index thread 0 takes from the queue head.
template <typename TS>
__device__ int64_t next_problem_block(const Args<TS>& a, int& round, int* slot) {
  int64_t idx;
  if (a.qhead) {
    __syncthreads();
    if (threadIdx.x == 0) *slot = (int)atomicAdd(a.qhead, 1u);
    __syncthreads();
    idx = *slot;
  } else {
    if (round++ > 0) return -1;
    idx = blockIdx.x;
  }
  const int64_t count = a.order ? a.order_count : a.B;
  if (idx >= count) return -1;
  return a.order ? (int64_t)a.order[idx] : idx;
}

// ------------------------------------------------------------------------------------------
// setup: load + sanitise + initialise (qpx/pdip.py:204-225)
// ------------------------------------------------------------------------------------------
template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_setup(Args<TS> a) {
  __shared__ double red[NT / 64];
  const PIdx pi = pidx(a, blockIdx.x);
  const View v = view_of(a, pi);
  stage_load<NT, TS>(v, a.Q + pi.oQ, a.q + pi.oq, a.A + pi.oA, a.b + pi.ob, a.G + pi.oG, a.h + pi.oh);
  stage_init<NT>(v, red);
}

// Active predicate of jax.lax.while_loop(iter < 30 and converged == 0) (pdip.py:228-232).
__device__ __forceinline__ bool is_active(const View& v, bool relax) {
  const int c = v.ctl[relax ? kRelaxConv : kConv], it = v.ctl[relax ? kRelaxIter : kIter];
  return it < kMaxIters && c == 0;
}

// ---------------------------------- plain-mode stage kernels --------------------------------
template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_resid(Args<TS> a, int relax) {
  __shared__ double red[NT / 64];
  if (blockIdx.x == 0 && threadIdx.x == 0) *a.counter = 0u;
  const View v = view_of(a, blockIdx.x);
  const bool act = is_active(v, relax);
  __syncthreads();
  if (threadIdx.x == 0) v.ctl[kActive] = act ? 1 : 0;
  if (!act) return;
  const int conv = stage_residuals<NT>(v, relax ? a.kappa : 0.0, a.tol, red);
  if (threadIdx.x == 0) v.ctl[5] = conv;  // entry flag, committed by k_step
}

template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_factor(Args<TS> a) {
  const View v = view_of(a, blockIdx.x);
  if (!v.ctl[kActive]) return;
  stage_factor<NT>(v, kDzs);
}

template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_solve(Args<TS> a, int affine) {
  const View v = view_of(a, blockIdx.x);
  if (!v.ctl[kActive]) return;
  if (affine) stage_solve<NT>(v, v.r1, v.r2, v.r3, v.r4, -1.0, v.dx, v.dsa, v.dza);
  else stage_solve<NT>(v, v.r1, v.r2, v.r3, v.r4, -1.0, v.dx, v.ds, v.dz);
}

template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_center(Args<TS> a) {
  __shared__ double red[NT / 64];
  const View v = view_of(a, blockIdx.x);
  if (!v.ctl[kActive]) return;
  stage_center<NT>(v, red);
}

template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_step(Args<TS> a, int relax) {
  __shared__ double red[NT / 64];
  const View v = view_of(a, blockIdx.x);
  if (!v.ctl[kActive]) return;
  stage_step<NT>(v, relax != 0, red);
  if (threadIdx.x == 0) {
    v.ctl[relax ? kRelaxConv : kConv] = v.ctl[5];
    v.ctl[relax ? kRelaxIter : kIter] += 1;
    if (is_active(v, relax)) atomicAdd(a.counter, 1u);
  }
}

// ------------------------- plain-mode stage kernels with LDS images (n < 64) -----------------
// For the sizes below the matrix-core path the plain stages work in LDS: each kernel stages the
// caller's G (read in its input precision with coalesced loads, rows with infinite h zeroed as
// pdip.py:176 does) and what else it reuses -- symmetrised Q and A for the residuals, the factor
// for the solves -- into the workgroup's LDS, runs the same stage code on LDS pointers (Cholesky
// of H, W = L^-1 A^T, F = A H^-1 A^T and the triangular sweeps among them), and writes back what
// the next stage reads.  Vectors move with them.
__host__ __device__ inline int64_t lds_words(int64_t k) { return round_up(k > 0 ? k : 1, 2); }
__host__ __device__ inline int plain_lds_bytes(int n, int me, int mi) {
  const int64_t G = lds_words((int64_t)mi * n), H = lds_words((int64_t)n * n), W = lds_words((int64_t)n * me),
                F = lds_words((int64_t)me * me), vn = lds_words(n), vm = lds_words(mi), ve = lds_words(me);
  const int64_t resid = G + H + lds_words((int64_t)me * n) + 3 * vn + 4 * vm + 3 * ve;  // G, Q, A; vectors
  const int64_t factor = G + H + W + F + 2 * vn + 2 * ve + vm;                          // + hdiag/hrcp, fdiag/frcp, D
  const int64_t solve = G + H + W + F + 2 * vn + 2 * ve + vm + 3 * vn + 7 * vm + 4 * ve;
  int64_t w = resid > factor ? resid : factor;
  w = w > solve ? w : solve;
  return (int)(8 * w);
}
// two workgroups per CU beside the stage code's static LDS (block reductions, the blocked sweeps'
// right-hand-side buffers).  n and m_eq below the matrix-core sizes: the tiled factorisation and
// sweeps (n or m_eq >= 64) read their factor through global-memory pointers.
constexpr int kPlainLdsMax = 74 * 1024;
__host__ __device__ inline bool plain_lds(int n, int me, int mi) {
  return n < kMfmaMinN && me < kMfmaMinN && plain_lds_bytes(n, me, mi) <= kPlainLdsMax;
}

struct LdsCarve {
  double* p;
  __device__ double* take(int64_t k) {
    double* r = p;
    p += lds_words(k);
    return r;
  }
};
template <int NT>
__device__ __forceinline__ void copy_vec(double* dst, const double* src, int count) {
  for (int e = threadIdx.x; e < count; e += NT) dst[e] = src[e];
}
// G from the caller's array into LDS (float64), rows with infinite h zeroed (0 * G, as stage_load)
template <int NT, typename TS>
__device__ __forceinline__ void stage_G_lds(double* dst, const TS* G, const double* hinf, int mi, int n) {
  for (int e = threadIdx.x; e < mi * n; e += NT) {
    const int k = e / n;
    const double g = (double)G[e];
    dst[e] = hinf[k] != 0.0 ? 0.0 * g : g;
  }
}

template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_resid_lds(Args<TS> a, int relax) {
  __shared__ double red[NT / 64];
  QPX_DYN_LDS(qpx_smem);
  if (blockIdx.x == 0 && threadIdx.x == 0) *a.counter = 0u;
  const PIdx pi = pidx(a, blockIdx.x);
  const View g = view_of(a, pi);
  const bool act = is_active(g, relax);
  __syncthreads();
  if (threadIdx.x == 0) g.ctl[kActive] = act ? 1 : 0;
  if (!act) return;
  const int n = pi.n, me = pi.me, mi = pi.mi;
  LdsCarve c{reinterpret_cast<double*>(qpx_smem)};
  View v = g;
  v.G = c.take((int64_t)mi * n);
  v.Q = c.take((int64_t)n * n);
  v.A = c.take((int64_t)me * n);
  v.x = c.take(n); v.q = c.take(n); v.r1 = c.take(n);
  v.s = c.take(mi); v.z = c.take(mi); v.h = c.take(mi); v.r3 = c.take(mi);
  v.y = c.take(me); v.b = c.take(me); v.r4 = c.take(me);
  v.r2 = g.r2;
  stage_G_lds<NT, TS>(v.G, a.G + pi.oG, g.hinf, mi, n);
  for (int e = threadIdx.x; e < n * n; e += NT) v.Q[e] = (double)a.Q[pi.oQ + e];  // symmetrised below
  for (int e = threadIdx.x; e < me * n; e += NT) v.A[e] = (double)a.A[pi.oA + e];
  copy_vec<NT>(v.x, g.x, n); copy_vec<NT>(v.q, g.q, n);
  copy_vec<NT>(v.s, g.s, mi); copy_vec<NT>(v.z, g.z, mi); copy_vec<NT>(v.h, g.h, mi);
  copy_vec<NT>(v.y, g.y, me); copy_vec<NT>(v.b, g.b, me);
  __syncthreads();
  for (int e = threadIdx.x; e < n * n; e += NT) {  // Q symmetrised as pdip.py:223 (same bits as stage_load)
    const int i = e / n, j = e - i * n;
    if (j < i) {
      const double sym = 0.5 * (v.Q[e] + v.Q[j * n + i]);
      v.Q[e] = sym;
      v.Q[j * n + i] = sym;
    }
  }
  __syncthreads();
  const int conv = stage_residuals<NT>(v, relax ? a.kappa : 0.0, a.tol, red);
  copy_vec<NT>(g.r1, v.r1, n);
  copy_vec<NT>(g.r3, v.r3, mi);
  copy_vec<NT>(g.r4, v.r4, me);
  if (threadIdx.x == 0) g.ctl[5] = conv;  // entry flag, committed by k_step
}

// all_problems: the backward's factorisation, which every problem takes
template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_factor_lds(Args<TS> a, int all_problems) {
  QPX_DYN_LDS(qpx_smem);
  const PIdx pi = pidx(a, blockIdx.x);
  const View g = view_of(a, pi);
  if (!all_problems && !g.ctl[kActive]) return;
  const int n = pi.n, me = pi.me, mi = pi.mi;
  LdsCarve c{reinterpret_cast<double*>(qpx_smem)};
  View v = g;
  v.G = c.take((int64_t)mi * n);
  v.H = c.take((int64_t)n * n);
  v.W = c.take((int64_t)n * me);
  v.F = c.take((int64_t)me * me);
  v.hdiag = c.take(n); v.hrcp = c.take(n); v.fdiag = c.take(me); v.frcp = c.take(me);
  v.D = c.take(mi);
  stage_G_lds<NT, TS>(v.G, a.G + pi.oG, g.hinf, mi, n);
  __syncthreads();
  stage_factor<NT>(v, kDzs);
  for (int e = threadIdx.x; e < n * n; e += NT) g.H[e] = v.H[e];
  copy_vec<NT>(g.W, v.W, n * me);
  copy_vec<NT>(g.F, v.F, me * me);
  copy_vec<NT>(g.hrcp, v.hrcp, n);
  copy_vec<NT>(g.frcp, v.frcp, me);
  copy_vec<NT>(g.D, v.D, mi);
}

// Solve stage on LDS: the right-hand side (a1..a4) and the outputs are the View's vectors named by
// `which` (0: affine direction into dsa/dza, 1: full direction into ds/dz; dx, dy either way).
template <int NT, typename TS>
__device__ void solve_lds(const Args<TS>& a, const PIdx& pi, const View& g, int which) {
  QPX_DYN_LDS(qpx_smem);
  const int n = pi.n, me = pi.me, mi = pi.mi;
  LdsCarve c{reinterpret_cast<double*>(qpx_smem)};
  View v = g;
  v.G = c.take((int64_t)mi * n);
  v.H = c.take((int64_t)n * n);
  v.W = c.take((int64_t)n * me);
  v.F = c.take((int64_t)me * me);
  v.hdiag = c.take(n); v.hrcp = c.take(n); v.fdiag = c.take(me); v.frcp = c.take(me);
  v.D = c.take(mi);
  v.r1 = c.take(n); v.u = c.take(n); v.dx = c.take(n);
  v.r2 = c.take(mi); v.r3 = c.take(mi); v.z = c.take(mi); v.s = c.take(mi);
  double* ds = c.take(mi);
  double* dz = c.take(mi);
  v.r4 = c.take(me); v.dy = c.take(me); v.t = c.take(me);
  stage_G_lds<NT, TS>(v.G, a.G + pi.oG, g.hinf, mi, n);
  for (int e = threadIdx.x; e < n * n; e += NT) v.H[e] = g.H[e];
  copy_vec<NT>(v.W, g.W, n * me);
  copy_vec<NT>(v.F, g.F, me * me);
  copy_vec<NT>(v.hrcp, g.hrcp, n);
  copy_vec<NT>(v.frcp, g.frcp, me);
  copy_vec<NT>(v.D, g.D, mi);
  copy_vec<NT>(v.r1, g.r1, n);
  copy_vec<NT>(v.r2, g.r2, mi); copy_vec<NT>(v.r3, g.r3, mi);
  copy_vec<NT>(v.z, g.z, mi); copy_vec<NT>(v.s, g.s, mi);
  copy_vec<NT>(v.r4, g.r4, me);
  __syncthreads();
  stage_solve<NT>(v, v.r1, v.r2, v.r3, v.r4, -1.0, v.dx, ds, dz);
  copy_vec<NT>(g.dx, v.dx, n);
  copy_vec<NT>(g.dy, v.dy, me);
  copy_vec<NT>(which ? g.ds : g.dsa, ds, mi);
  copy_vec<NT>(which ? g.dz : g.dza, dz, mi);
}

template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_solve_lds(Args<TS> a, int affine) {
  const PIdx pi = pidx(a, blockIdx.x);
  const View g = view_of(a, pi);
  if (!g.ctl[kActive]) return;
  solve_lds<NT, TS>(a, pi, g, affine ? 0 : 1);
}

template <int NT, typename TS>
__device__ void bwd_write(const Args<TS>& a, const View& v, int64_t p);

// The backward's solve (right-hand side (-dl/dx, 0, 0, 0) from bwd_load) on LDS, then the
// cotangents from the directions it wrote back.
template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_bwd_solve_lds(Args<TS> a) {
  const PIdx pi = pidx(a, blockIdx.x);
  const View g = view_of(a, pi);
  solve_lds<NT, TS>(a, pi, g, 1);
  __syncthreads();
  bwd_write<NT, TS>(a, g, blockIdx.x);
}

// Post-processing of solve_qp (pdip.py:243-251) and start of the relaxation.
template <int NT, typename TS>
__device__ void finish_solve(const Args<TS>& a, const View& v, int64_t p) {
  const PIdx pi = pidx(a, p);
  const int n = pi.n, me = pi.me, mi = pi.mi, tid = threadIdx.x;
  for (int k = tid; k < mi; k += NT) {
    if (v.hinf[k] != 0.0) { v.z[k] = 0.0; v.s[k] = INFINITY; }
    a.s[pi.oh + k] = (TS)v.s[k];
    a.z[pi.oh + k] = (TS)v.z[k];
  }
  for (int i = tid; i < n; i += NT) a.x[pi.oq + i] = (TS)v.x[i];
  for (int c = tid; c < me; c += NT) a.y[pi.ob + c] = (TS)v.y[c];
  __syncthreads();
  if (tid == 0) {
    a.conv[p] = v.ctl[kConv];
    a.iters[p] = v.ctl[kIter];
    v.ctl[kRelaxConv] = 0;
    v.ctl[kRelaxIter] = 0;
  }
  __syncthreads();
}

template <int NT, typename TS>
__device__ void finish_relax(const Args<TS>& a, const View& v, int64_t p) {
  const PIdx pi = pidx(a, p);
  const int n = pi.n, me = pi.me, mi = pi.mi, tid = threadIdx.x;
  for (int k = tid; k < mi; k += NT) { a.sr[pi.oh + k] = (TS)v.s[k]; a.zr[pi.oh + k] = (TS)v.z[k]; }
  for (int i = tid; i < n; i += NT) a.xr[pi.oq + i] = (TS)v.x[i];
  for (int c = tid; c < me; c += NT) a.yr[pi.ob + c] = (TS)v.y[c];
  if (tid == 0) { a.rconv[p] = v.ctl[kRelaxConv]; a.riters[p] = v.ctl[kRelaxIter]; }
}

template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_finish_solve(Args<TS> a) {
  finish_solve<NT, TS>(a, view_of(a, blockIdx.x), blockIdx.x);
}

template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_finish_relax(Args<TS> a) {
  finish_relax<NT, TS>(a, view_of(a, blockIdx.x), blockIdx.x);
}

// ------------------------------------ fused forward ----------------------------------------
template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_fused_solve(Args<TS> a, int with_relax) {
  __shared__ double red[NT / 64];
  __shared__ int slot;
  int round = 0;
  for (int64_t p; (p = next_problem_block(a, round, &slot)) >= 0;) {
    const PIdx pi = pidx(a, p);
    const View v = view_of(a, pi);
    stage_load<NT, TS>(v, a.Q + pi.oQ, a.q + pi.oq, a.A + pi.oA, a.b + pi.ob, a.G + pi.oG, a.h + pi.oh);
    stage_init<NT>(v, red);
    double* tr = a.trace ? a.trace + p * (int64_t)(kMaxIters * kTraceCols) : nullptr;
    while (is_active(v, false)) {
      __syncthreads();
      pc_iteration<NT>(v, a.tol, red, tr ? tr + v.ctl[kIter] * kTraceCols : nullptr, (a.flags & 4) != 0);
    }
    if (tr) {  // rows of iterations not taken
      __syncthreads();
      for (int e = v.ctl[kIter] * kTraceCols + threadIdx.x; e < kMaxIters * kTraceCols; e += NT) tr[e] = NAN;
    }
    finish_solve<NT, TS>(a, v, p);
    if (!with_relax) continue;
    while (is_active(v, true)) {
      __syncthreads();
      relax_iteration<NT>(v, a.tol, a.kappa, red);
    }
    finish_relax<NT, TS>(a, v, p);
  }
}

// initialize (qpax/pdip.py:15-35) alone: the load of solve_qp (Q symmetrised, rows with infinite
// h zeroed, pdip.py:220-223) and the CVXOPT initialisation, written to x, s, z, y.
template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_initialize(Args<TS> a) {
  __shared__ double red[NT / 64];
  const int64_t p = blockIdx.x;
  const PIdx pi = pidx(a, p);
  const View v = view_of(a, pi);
  stage_load<NT, TS>(v, a.Q + pi.oQ, a.q + pi.oq, a.A + pi.oA, a.b + pi.ob, a.G + pi.oG, a.h + pi.oh);
  stage_init<NT>(v, red);
  const int tid = threadIdx.x;
  for (int i = tid; i < pi.n; i += NT) a.x[pi.oq + i] = (TS)v.x[i];
  for (int k = tid; k < pi.mi; k += NT) { a.s[pi.oh + k] = (TS)v.s[k]; a.z[pi.oh + k] = (TS)v.z[k]; }
  for (int c = tid; c < pi.me; c += NT) a.y[pi.ob + c] = (TS)v.y[c];
}

// ------------------------------------ backward ---------------------------------------------
// diff_qp (qpax/diff_qp.py:20-46): load the relaxed point and the cotangent.
template <int NT, typename TS>
__device__ void bwd_load(const Args<TS>& a, const View& v, int64_t p) {
  const PIdx pi = pidx(a, p);
  const int n = pi.n, me = pi.me, mi = pi.mi;
  stage_load<NT, TS>(v, a.Q + pi.oQ, a.q + pi.oq, a.A + pi.oA, a.b + pi.ob, a.G + pi.oG, a.h + pi.oh);
  const int tid = threadIdx.x;
  for (int i = tid; i < n; i += NT) { v.x[i] = (double)a.xr_in[pi.oq + i]; v.r1[i] = (double)a.g[pi.oq + i]; }
  for (int k = tid; k < mi; k += NT) {
    v.s[k] = (double)a.sr_in[pi.oh + k];
    v.z[k] = (double)a.zr_in[pi.oh + k];
    v.r2[k] = 0.0;
    v.r3[k] = 0.0;
  }
  for (int c = tid; c < me; c += NT) { v.y[c] = (double)a.yr_in[pi.ob + c]; v.r4[c] = 0.0; }
  __syncthreads();
}

// Solve with rhs (-dl_dx, 0, 0, 0) and form the cotangents (optnet_derivatives, diff_qp.py:8-17).
template <int NT, typename TS>
__device__ void bwd_write(const Args<TS>& a, const View& v, int64_t p) {
  const PIdx pi = pidx(a, p);
  const int n = pi.n, me = pi.me, mi = pi.mi, tid = threadIdx.x;
  // dz here is dlam_tilde; dlam = dlam_tilde / lam
  for (int e = tid; e < n * n; e += NT) {
    const int i = e / n, j = e - i * n;
    a.dQ[pi.oQ + e] = (TS)(0.5 * (v.dx[i] * v.x[j] + v.x[i] * v.dx[j]));
  }
  for (int i = tid; i < n; i += NT) a.dq[pi.oq + i] = (TS)v.dx[i];
  for (int e = tid; e < me * n; e += NT) {
    const int c = e / n, j = e - c * n;
    a.dA[pi.oA + e] = (TS)(v.dy[c] * v.x[j] + v.y[c] * v.dx[j]);
  }
  for (int c = tid; c < me; c += NT) a.db[pi.ob + c] = (TS)(-v.dy[c]);
  for (int e = tid; e < mi * n; e += NT) {
    const int k = e / n, j = e - k * n;
    const double lam = v.z[k], dlam = v.dz[k] / lam;
    // reference: diag(lam) (dlam x^T + lam dx^T) (diff_qp.py:11); exact VJP: diag(lam) dlam x^T + lam dx^T
    const double val = (a.flags & 1) ? lam * dlam * v.x[j] + lam * v.dx[j] : lam * (dlam * v.x[j] + lam * v.dx[j]);
    a.dG[pi.oG + e] = (TS)val;
  }
  for (int k = tid; k < mi; k += NT) {
    const double lam = v.z[k], dlam = v.dz[k] / lam;
    a.dh[pi.oh + k] = (TS)(-lam * dlam);
  }
}

template <int NT, typename TS>
__device__ void bwd_solve_write(const Args<TS>& a, const View& v, int64_t p) {
  stage_solve<NT>(v, v.r1, v.r2, v.r3, v.r4, -1.0, v.dx, v.ds, v.dz);
  bwd_write<NT, TS>(a, v, p);
}

template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_bwd_load(Args<TS> a) {
  bwd_load<NT, TS>(a, view_of(a, blockIdx.x), blockIdx.x);
}

template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_bwd_factor(Args<TS> a) {
  stage_factor<NT>(view_of(a, blockIdx.x), kDzs);
}

template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_bwd_solve(Args<TS> a) {
  bwd_solve_write<NT, TS>(a, view_of(a, blockIdx.x), blockIdx.x);
}

template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_fused_bwd(Args<TS> a) {
  __shared__ int slot;
  int round = 0;
  for (int64_t p; (p = next_problem_block(a, round, &slot)) >= 0;) {
    const View v = view_of(a, p);
    bwd_load<NT, TS>(a, v, p);
    stage_factor<NT>(v, kDzs);
    bwd_solve_write<NT, TS>(a, v, p);
    __syncthreads();
  }
}


// ------------------------------ static special branches --------------------------------------
// qpax/pdip.py:208-217: with no inequality rows the reference skips the interior-point method.
// No constraints: x = solve(Q, -q) (jnp.linalg.solve, LU with partial pivoting) on the
// unsymmetrised Q.  Equality rows only: solve_eq_only (pdip.py:137-150, Boyd p.559),
// x = Q^{-1}(-A^T y - q) with y = S^{-1}(A Q^{-1} q + b), S = -A Q^{-1} A^T, also on the
// unsymmetrised Q.  Both factor with partial pivoting (the reference uses LU resp. QR).

// In-place LU with partial pivoting of M (m x m, row-major); perm[i] = original row of row i.
template <int NT>
__device__ void block_lu(double* M, int m, int* perm, double* red) {
  const int tid = threadIdx.x;
  for (int i = tid; i < m; i += NT) perm[i] = i;
  __syncthreads();
  for (int j = 0; j < m; ++j) {
    // pivot: first row of maximal |M[i][j]|, i >= j
    double best = -1.0;
    int bi = j;
    for (int i = j + tid; i < m; i += NT) {
      const double a = fabs(M[i * m + j]);
      if (a > best) { best = a; bi = i; }
    }
    const double bmax = block_reduce<NT>(best, red, OpMax());
    int cand = (best == bmax) ? bi : m;
    // lowest row index attaining the maximum
    for (int o = 32; o > 0; o >>= 1) cand = min(cand, __shfl_xor(cand, o, 64));
    if constexpr (NT > 64) {
      __syncthreads();
      if ((tid & 63) == 0) reinterpret_cast<int*>(red)[tid >> 6] = cand;
      __syncthreads();
      for (int w = 0; w < NT / 64; ++w) cand = min(cand, reinterpret_cast<int*>(red)[w]);
    }
    const int piv = cand < m ? cand : j;
    __syncthreads();
    if (piv != j) {
      for (int c = tid; c < m; c += NT) {
        const double t = M[j * m + c];
        M[j * m + c] = M[piv * m + c];
        M[piv * m + c] = t;
      }
      if (tid == 0) { const int t = perm[j]; perm[j] = perm[piv]; perm[piv] = t; }
    }
    __syncthreads();
    const double inv = 1.0 / M[j * m + j];
    for (int i = j + 1 + tid; i < m; i += NT) M[i * m + j] *= inv;
    __syncthreads();
    const int r = m - j - 1;
    for (int e = tid; e < r * r; e += NT) {
      const int a = e / r, c = e - a * r;
      const int i = j + 1 + a, k = j + 1 + c;
      M[i * m + k] -= M[i * m + j] * M[j * m + k];
    }
    __syncthreads();
  }
}

// Solve (LU) X = P B for `nrhs` right-hand sides stored column-major in B (m x nrhs, column c at
// B + c*m); results overwrite B.  tmp holds m * nrhs doubles.
template <int NT>
__device__ void block_lu_solve(const double* LU, const int* perm, int m, double* B, int nrhs, double* tmp) {
  const int tid = threadIdx.x;
  for (int e = tid; e < m * nrhs; e += NT) {
    const int c = e / m, i = e - c * m;
    tmp[e] = B[c * m + perm[i]];
  }
  __syncthreads();
  for (int j = 0; j < m; ++j) {  // unit lower
    for (int e = tid; e < (m - j - 1) * nrhs; e += NT) {
      const int c = e / (m - j - 1), i = j + 1 + (e - c * (m - j - 1));
      tmp[c * m + i] -= LU[i * m + j] * tmp[c * m + j];
    }
    __syncthreads();
  }
  for (int j = m - 1; j >= 0; --j) {  // upper
    for (int c = tid; c < nrhs; c += NT) tmp[c * m + j] /= LU[j * m + j];
    __syncthreads();
    for (int e = tid; e < j * nrhs; e += NT) {
      const int c = e / j, i = e - c * j;
      tmp[c * m + i] -= LU[i * m + j] * tmp[c * m + j];
    }
    __syncthreads();
  }
  for (int e = tid; e < m * nrhs; e += NT) B[e] = tmp[e];
  __syncthreads();
}

template <int NT, typename TS>
__global__ __launch_bounds__(NT) void k_special(Args<TS> a) {
  __shared__ double red[NT / 64];
  const int64_t p = blockIdx.x;
  const int n = a.n, me = a.me, tid = threadIdx.x;
  const View v = view_of(a, p);
  // workspace reuse: H (n x n) the LU of Q, G-region for right-hand sides, F the LU of S
  double* LU = v.H;
  int* perm = reinterpret_cast<int*>(v.hdiag);
  int* permS = reinterpret_cast<int*>(v.fdiag);
  double* rhs = v.G;                 // (me + 1) columns of length n
  double* tmp = v.G + (int64_t)(me + 1) * n;
  for (int e = tid; e < n * n; e += NT) LU[e] = (double)a.Q[p * n * n + e];
  for (int i = tid; i < n; i += NT) rhs[me * n + i] = (double)a.q[p * n + i];
  for (int e = tid; e < me * n; e += NT) {
    const int c = e / n, i = e - c * n;
    rhs[c * n + i] = (double)a.A[(p * me + c) * n + i];  // column c = A^T e_c
  }
  __syncthreads();
  block_lu<NT>(LU, n, perm, red);
  block_lu_solve<NT>(LU, perm, n, rhs, me + 1, tmp);  // rhs <- [Q^{-1} A^T, Q^{-1} q]
  const double* Xq = rhs + me * n;
  if (me == 0) {
    for (int i = tid; i < n; i += NT) a.x[p * n + i] = (TS)(-Xq[i]);
  } else {
    double* S = v.F;
    double* ys = v.frcp;   // me
    double* ytmp = v.t;    // me
    for (int e = tid; e < me * me; e += NT) {
      const int r = e / me, c = e - r * me;
      double acc = 0.0;
      for (int i = 0; i < n; ++i) acc += (double)a.A[(p * me + r) * n + i] * rhs[c * n + i];
      S[e] = -acc;
    }
    for (int r = tid; r < me; r += NT) {
      double acc = (double)a.b[p * me + r];
      for (int i = 0; i < n; ++i) acc += (double)a.A[(p * me + r) * n + i] * Xq[i];
      ys[r] = acc;
    }
    __syncthreads();
    block_lu<NT>(S, me, permS, red);
    block_lu_solve<NT>(S, permS, me, ys, 1, ytmp);
    for (int i = tid; i < n; i += NT) {
      double acc = -Xq[i];
      for (int c = 0; c < me; ++c) acc -= rhs[c * n + i] * ys[c];
      a.x[p * n + i] = (TS)acc;
    }
    for (int c = tid; c < me; c += NT) a.y[p * me + c] = (TS)ys[c];
  }
  if (tid == 0) { a.conv[p] = 1; a.iters[p] = 0; }
}

}  // namespace qpx

// ===================================== host side =============================================
namespace qpx {

// Launchers of the workgroup-per-QP fused kernels, shared with the ragged-batch scheduler.
template <int NT, typename TS>
hipError_t launch_generic_solve(const Args<TS>& a, unsigned grid, int lds, hipStream_t st, int relax) {
  hipLaunchKernelGGL((k_fused_solve<NT, TS>), dim3(grid), dim3(NT), lds, st, a, relax);
  return hipGetLastError();
}

template <int NT, typename TS>
hipError_t launch_generic_bwd(const Args<TS>& a, unsigned grid, int lds, hipStream_t st) {
  hipLaunchKernelGGL((k_fused_bwd<NT, TS>), dim3(grid), dim3(NT), lds, st, a);
  return hipGetLastError();
}

template <int NT, typename TS>
int generic_occupancy(int lds) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(&k_fused_solve<NT, TS>), NT,
                                                   (size_t)lds) != hipSuccess)
    nb = 0;
  return nb;
}

#define QPX_GEN_INST(NT, TS)                                                                        \
  template hipError_t launch_generic_solve<NT, TS>(const Args<TS>&, unsigned, int, hipStream_t, int); \
  template hipError_t launch_generic_bwd<NT, TS>(const Args<TS>&, unsigned, int, hipStream_t);        \
  template int generic_occupancy<NT, TS>(int);
QPX_GEN_INST(64, float)
QPX_GEN_INST(128, float)
QPX_GEN_INST(256, float)
QPX_GEN_INST(64, double)
QPX_GEN_INST(128, double)
QPX_GEN_INST(256, double)
#undef QPX_GEN_INST

namespace host {

namespace {
thread_local std::string g_err;
}

qpx_status fail(qpx_status st, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return st;
}

void clear_error() { g_err.clear(); }
const char* last_error() { return g_err.c_str(); }

// Per-problem float64 words: the generic kernels' view, or the wavefront kernels' m x m reduced
// KKT matrix (m = n padded to a multiple of 4, plus m_eq), whichever is larger.
int64_t stride_words(int n, int me, int mi) {
  const int64_t m = qpx::wave::padded_n(n) + me;
  int64_t w = qpx::elastic_words(n, me, mi);
  if (w < qpx::round_up(m * m, 64)) w = qpx::round_up(m * m, 64);
  if (w < qpx::wave::split_words(n, me)) w = qpx::wave::split_words(n, me);
  const int64_t g = qpx::group::group_words(n, me, mi);
  return w > g ? w : g;
}

int max_lds_per_block() {
  static int v = -1;
  if (v < 0) {
    int dev = 0;
    int val = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&val, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) val = 65536;
    v = val;
  }
  return v;
}

int split_spin_limit(int dflt) {
  const char* v = std::getenv("QPX_SPLIT_SPIN");
  if (!v || !*v) return dflt;
  const long r = std::strtol(v, nullptr, 10);
  return r < 0 ? 0 : (r > dflt ? dflt : (int)r);
}

int cu_count() {
  int dev = 0, val = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&val, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    val = 256;
  return val;
}

// The wavefront-per-QP kernels hold the augmented KKT row of a lane in registers, so they take
// problems with n + m_eq <= 64 and m_ineq <= 256 whose LDS image fits one workgroup.
bool wave_variant(int n, int me, int mi, int ts, WaveVariant* v) {
  const int m = qpx::wave::padded_n(n) + me;
  if (m > 64 || mi > 256) return false;
  if (qpx::wave::lds_layout(n, me, mi, ts).bytes > max_lds_per_block()) return false;
  if (m <= 32 && mi <= 64) *v = {32, 1};
  else if (m <= 32 && mi <= 128) *v = {32, 2};
  else if (mi <= 64) *v = {64, 1};
  else if (mi <= 128) *v = {64, 2};
  else *v = {64, 4};
  return true;
}

bool group_variant(int n, int me, int mi, int ts, int w, GroupVariant* v) {
  const int m = qpx::wave::padded_n(n) + me;
  int nti = qpx::group::ntiles(m);
  if (nti > 10 || mi > 512) return false;
  if (nti < 4) nti = 4;
  if (w != 4 && w != 8) w = 4;
  *v = {nti, w};
  return group_lds_bytes(n, me, mi, ts, *v) + 256 <= max_lds_per_block();  // + the kernel's static LDS
}

int group_lds_bytes(int n, int me, int mi, int ts, GroupVariant v) {
  const int m = qpx::wave::padded_n(n) + me;
  (void)v.nti;
  return qpx::group::group_lds(m, mi, 64 * v.w, ts).bytes;
}

}  // namespace host
}  // namespace qpx

namespace {
using qpx::host::fail;
using qpx::host::kCtlBytes;
using qpx::host::stride_words;

size_t ws_bytes(int64_t B, int n, int me, int mi) {
  return kCtlBytes + (size_t)B * (size_t)stride_words(n, me, mi) * sizeof(double);
}

qpx_status check_dims(int32_t dtype, int32_t mode, int64_t B, int32_t n, int32_t me, int32_t mi,
                      size_t have) {
  if (dtype != QPX_F32 && dtype != QPX_F64) return fail(QPX_ERR_ARG, "dtype must be QPX_F32 or QPX_F64 (got %d)", dtype);
  if (mode != QPX_PLAIN && mode != QPX_FUSED) return fail(QPX_ERR_ARG, "mode must be QPX_PLAIN or QPX_FUSED (got %d)", mode);
  if (B < 0 || n < 1 || me < 0 || mi < 0) return fail(QPX_ERR_ARG, "invalid sizes batch=%lld n=%d m_eq=%d m_ineq=%d", (long long)B, n, me, mi);
  if (mi < 1) return fail(QPX_ERR_ARG, "m_ineq must be >= 1 for the interior-point path (got %d)", mi);
  if (B > 0x7fffffffLL) return fail(QPX_ERR_UNSUPPORTED, "batch %lld exceeds 2^31-1", (long long)B);
  const size_t need = ws_bytes(B, n, me, mi);
  if (have < need) return fail(QPX_ERR_WORKSPACE, "workspace of %zu bytes is smaller than the %zu required", have, need);
  return QPX_OK;
}

template <typename TS>
qpx::Args<TS> make_args(int64_t B, int n, int me, int mi, void* ws) {
  qpx::Args<TS> a;
  std::memset(&a, 0, sizeof a);
  a.B = B; a.n = n; a.me = me; a.mi = mi;
  a.counter = reinterpret_cast<unsigned*>(ws);
  a.ws = reinterpret_cast<double*>(static_cast<char*>(ws) + kCtlBytes);
  a.stride = stride_words(n, me, mi);
  return a;
}

int* pinned_counter() {
  thread_local int* p = nullptr;
  if (!p) {
    if (hipHostMalloc(reinterpret_cast<void**>(&p), sizeof(int), hipHostMallocDefault) != hipSuccess) p = nullptr;
  }
  return p;
}

// A uniform batch on the wavefront kernels: one wavefront (workgroup) per problem.
template <typename TS>
bool wave_plan(qpx::Args<TS>& a, qpx::host::WaveVariant* v) {
  if (!qpx::host::wave_variant(a.n, a.me, a.mi, (int)sizeof(TS), v)) return false;
  a.lds_per_wave = qpx::wave::lds_layout(a.n, a.me, a.mi, (int)sizeof(TS)).bytes;
  return true;
}

template <int NT, typename TS>
qpx_status run_forward(qpx::Args<TS>& a, int mode, bool relax, hipStream_t st) {
  const int lds = qpx::generic_lds_bytes(a.n, a.me, NT);  // G^T D G staging, panels, W (n >= 64)
  const dim3 grid((unsigned)a.B), block(NT);
  if (a.B == 0) return QPX_OK;
  if (mode == QPX_FUSED) {
    qpx::host::WaveVariant wv;
    if (wave_plan(a, &wv)) {
      QPX_HIP(qpx::wave::launch_solve(wv, a, (unsigned)a.B, 1, st, relax ? 1 : 0));
      return QPX_OK;
    }
    qpx::host::GroupVariant gv;
    if (qpx::host::group_variant(a.n, a.me, a.mi, (int)sizeof(TS), 0, &gv)) {
      const int lds = qpx::host::group_lds_bytes(a.n, a.me, a.mi, (int)sizeof(TS), gv);
      QPX_HIP(qpx::group::launch_solve(gv, a, (unsigned)a.B, lds, st, relax ? 1 : 0));
      return QPX_OK;
    }
    hipLaunchKernelGGL((qpx::k_fused_solve<NT, TS>), grid, block, lds, st, a, relax ? 1 : 0);
    QPX_HIP(hipGetLastError());
    return QPX_OK;
  }
  int* host = pinned_counter();
  if (!host) return fail(QPX_ERR_HIP, "hipHostMalloc of the plain-mode counter failed");
  hipLaunchKernelGGL((qpx::k_setup<NT, TS>), grid, block, lds, st, a);
  QPX_HIP(hipGetLastError());
  const bool in_lds = qpx::plain_lds(a.n, a.me, a.mi);
  const int plds = qpx::plain_lds_bytes(a.n, a.me, a.mi);
  for (int phase = 0; phase < (relax ? 2 : 1); ++phase) {
    const int rl = phase;
    for (int it = 0; it < qpx::kMaxIters; ++it) {
      if (in_lds) {
        hipLaunchKernelGGL((qpx::k_resid_lds<NT, TS>), grid, block, plds, st, a, rl);
        hipLaunchKernelGGL((qpx::k_factor_lds<NT, TS>), grid, block, plds, st, a, 0);
        if (!rl) {
          hipLaunchKernelGGL((qpx::k_solve_lds<NT, TS>), grid, block, plds, st, a, 1);
          hipLaunchKernelGGL((qpx::k_center<NT, TS>), grid, block, 0, st, a);
        }
        hipLaunchKernelGGL((qpx::k_solve_lds<NT, TS>), grid, block, plds, st, a, 0);
        hipLaunchKernelGGL((qpx::k_step<NT, TS>), grid, block, 0, st, a, rl);
      } else {
        hipLaunchKernelGGL((qpx::k_resid<NT, TS>), grid, block, lds, st, a, rl);
        hipLaunchKernelGGL((qpx::k_factor<NT, TS>), grid, block, lds, st, a);
        if (!rl) {
          hipLaunchKernelGGL((qpx::k_solve<NT, TS>), grid, block, lds, st, a, 1);
          hipLaunchKernelGGL((qpx::k_center<NT, TS>), grid, block, lds, st, a);
        }
        hipLaunchKernelGGL((qpx::k_solve<NT, TS>), grid, block, lds, st, a, 0);
        hipLaunchKernelGGL((qpx::k_step<NT, TS>), grid, block, lds, st, a, rl);
      }
      QPX_HIP(hipGetLastError());
      QPX_HIP(hipMemcpyAsync(host, a.counter, sizeof(int), hipMemcpyDeviceToHost, st));
      QPX_HIP(hipStreamSynchronize(st));
      if (*host == 0) break;
    }
    if (!rl) hipLaunchKernelGGL((qpx::k_finish_solve<NT, TS>), grid, block, lds, st, a);
    else hipLaunchKernelGGL((qpx::k_finish_relax<NT, TS>), grid, block, lds, st, a);
    QPX_HIP(hipGetLastError());
  }
  return QPX_OK;
}

template <int NT, typename TS>
qpx_status run_backward(qpx::Args<TS>& a, int mode, hipStream_t st) {
  const int lds = qpx::generic_lds_bytes(a.n, a.me, NT);  // G^T D G staging, panels, W (n >= 64)
  const dim3 grid((unsigned)a.B), block(NT);
  if (a.B == 0) return QPX_OK;
  if (mode == QPX_FUSED) {
    qpx::host::WaveVariant wv;
    if (wave_plan(a, &wv)) {
      QPX_HIP(qpx::wave::launch_bwd(wv, a, (unsigned)a.B, 1, st));
      return QPX_OK;
    }
    qpx::host::GroupVariant gv;
    if (qpx::host::group_variant(a.n, a.me, a.mi, (int)sizeof(TS), 0, &gv)) {
      const int lds = qpx::host::group_lds_bytes(a.n, a.me, a.mi, (int)sizeof(TS), gv);
      QPX_HIP(qpx::group::launch_bwd(gv, a, (unsigned)a.B, lds, st));
      return QPX_OK;
    }
    hipLaunchKernelGGL((qpx::k_fused_bwd<NT, TS>), grid, block, lds, st, a);
  } else if (qpx::plain_lds(a.n, a.me, a.mi)) {
    const int plds = qpx::plain_lds_bytes(a.n, a.me, a.mi);
    hipLaunchKernelGGL((qpx::k_bwd_load<NT, TS>), grid, block, lds, st, a);
    hipLaunchKernelGGL((qpx::k_factor_lds<NT, TS>), grid, block, plds, st, a, 1);
    hipLaunchKernelGGL((qpx::k_bwd_solve_lds<NT, TS>), grid, block, plds, st, a);
  } else {
    hipLaunchKernelGGL((qpx::k_bwd_load<NT, TS>), grid, block, lds, st, a);
    hipLaunchKernelGGL((qpx::k_bwd_factor<NT, TS>), grid, block, lds, st, a);
    hipLaunchKernelGGL((qpx::k_bwd_solve<NT, TS>), grid, block, lds, st, a);
  }
  QPX_HIP(hipGetLastError());
  return QPX_OK;
}

// Threads per block by problem size: one wavefront for small problems, more for large ones.
template <typename F>
qpx_status dispatch_nt(int n, int mi, F&& f) {
  if (n <= 32 && mi <= 128) return f(std::integral_constant<int, 64>());
  if (n <= 96) return f(std::integral_constant<int, 128>());
  return f(std::integral_constant<int, 256>());
}

template <typename TS>
qpx_status forward_impl(int mode, int64_t B, int n, int me, int mi, const void* Q, const void* q,
                        const void* A, const void* b, const void* G, const void* h, double tol,
                        double kappa, bool relax, void* x, void* s, void* z, void* y, int* conv,
                        int* iters, void* xr, void* sr, void* zr, void* yr, int* rconv,
                        int* riters, void* ws, void* stream, int64_t stride = 0) {
  auto a = make_args<TS>(B, n, me, mi, ws);
  if (stride > 0) a.stride = stride;
  a.Q = (const TS*)Q; a.q = (const TS*)q; a.A = (const TS*)A; a.b = (const TS*)b;
  a.G = (const TS*)G; a.h = (const TS*)h;
  a.x = (TS*)x; a.s = (TS*)s; a.z = (TS*)z; a.y = (TS*)y; a.conv = conv; a.iters = iters;
  a.xr = (TS*)xr; a.sr = (TS*)sr; a.zr = (TS*)zr; a.yr = (TS*)yr; a.rconv = rconv; a.riters = riters;
  a.tol = tol; a.kappa = kappa;
  return dispatch_nt(n, mi, [&](auto nt) { return run_forward<decltype(nt)::value, TS>(a, mode, relax, (hipStream_t)stream); });
}

template <typename TS>
qpx_status backward_impl(int mode, int64_t B, int n, int me, int mi, const void* Q, const void* q,
                         const void* A, const void* b, const void* G, const void* h,
                         const void* xr, const void* sr, const void* zr, const void* yr,
                         const void* g, int flags, void* dQ, void* dq, void* dA, void* db, void* dG, void* dh,
                         void* ws, void* stream) {
  auto a = make_args<TS>(B, n, me, mi, ws);
  a.flags = flags;
  a.Q = (const TS*)Q; a.q = (const TS*)q; a.A = (const TS*)A; a.b = (const TS*)b;
  a.G = (const TS*)G; a.h = (const TS*)h;
  a.xr_in = (const TS*)xr; a.sr_in = (const TS*)sr; a.zr_in = (const TS*)zr; a.yr_in = (const TS*)yr;
  a.g = (const TS*)g;
  a.dQ = (TS*)dQ; a.dq = (TS*)dq; a.dA = (TS*)dA; a.db = (TS*)db; a.dG = (TS*)dG; a.dh = (TS*)dh;
  return dispatch_nt(n, mi, [&](auto nt) { return run_backward<decltype(nt)::value, TS>(a, mode, (hipStream_t)stream); });
}

// ------------------------------------ elastic -----------------------------------------------
template <int NT, typename TS>
qpx_status run_e_forward(qpx::Args<TS>& a, int mode, bool relax, hipStream_t st) {
  const int lds = qpx::generic_lds_bytes(a.n, a.me, NT);  // G^T D G staging, panels, W (n >= 64)
  const dim3 grid((unsigned)a.B), block(NT);
  if (a.B == 0) return QPX_OK;
  if (mode == QPX_FUSED) {
    hipLaunchKernelGGL((qpx::k_e_fused_solve<NT, TS>), grid, block, lds, st, a, relax ? 1 : 0);
    QPX_HIP(hipGetLastError());
    return QPX_OK;
  }
  int* host = pinned_counter();
  if (!host) return fail(QPX_ERR_HIP, "hipHostMalloc of the plain-mode counter failed");
  hipLaunchKernelGGL((qpx::k_e_setup<NT, TS>), grid, block, lds, st, a);
  QPX_HIP(hipGetLastError());
  for (int phase = 0; phase < (relax ? 2 : 1); ++phase) {
    const int rl = phase;
    for (int it = 0; it < qpx::kMaxIters; ++it) {
      hipLaunchKernelGGL((qpx::k_e_resid<NT, TS>), grid, block, lds, st, a, rl);
      hipLaunchKernelGGL((qpx::k_e_factor<NT, TS>), grid, block, lds, st, a);
      if (!rl) {
        hipLaunchKernelGGL((qpx::k_e_solve<NT, TS>), grid, block, lds, st, a, 1);
        hipLaunchKernelGGL((qpx::k_e_center<NT, TS>), grid, block, lds, st, a);
      }
      hipLaunchKernelGGL((qpx::k_e_solve<NT, TS>), grid, block, lds, st, a, 0);
      hipLaunchKernelGGL((qpx::k_e_step<NT, TS>), grid, block, lds, st, a, rl);
      QPX_HIP(hipGetLastError());
      QPX_HIP(hipMemcpyAsync(host, a.counter, sizeof(int), hipMemcpyDeviceToHost, st));
      QPX_HIP(hipStreamSynchronize(st));
      if (*host == 0) break;
    }
    if (!rl) hipLaunchKernelGGL((qpx::k_e_finish_solve<NT, TS>), grid, block, lds, st, a);
    else hipLaunchKernelGGL((qpx::k_e_finish_relax<NT, TS>), grid, block, lds, st, a);
    QPX_HIP(hipGetLastError());
  }
  return QPX_OK;
}

template <int NT, typename TS>
qpx_status run_e_backward(qpx::Args<TS>& a, int mode, hipStream_t st) {
  const int lds = qpx::generic_lds_bytes(a.n, a.me, NT);  // G^T D G staging, panels, W (n >= 64)
  const dim3 grid((unsigned)a.B), block(NT);
  if (a.B == 0) return QPX_OK;
  if (mode == QPX_FUSED) {
    hipLaunchKernelGGL((qpx::k_e_fused_bwd<NT, TS>), grid, block, lds, st, a);
  } else {
    hipLaunchKernelGGL((qpx::k_e_bwd_load<NT, TS>), grid, block, lds, st, a);
    hipLaunchKernelGGL((qpx::k_e_bwd_factor<NT, TS>), grid, block, lds, st, a);
    hipLaunchKernelGGL((qpx::k_e_bwd_solve<NT, TS>), grid, block, lds, st, a);
  }
  QPX_HIP(hipGetLastError());
  return QPX_OK;
}

template <typename TS>
qpx_status e_forward_impl(int mode, int64_t B, int n, int me, int mi, const void* Q, const void* q, const void* A,
                          const void* b, const void* G, const void* h, const void* pen, double tol, double kappa,
                          int flags, bool relax, void* const* out, int* conv, int* iters, void* const* rout,
                          int* rconv, int* riters, void* ws, void* stream) {
  auto a = make_args<TS>(B, n, me, mi, ws);
  a.Q = (const TS*)Q; a.q = (const TS*)q; a.A = (const TS*)A; a.b = (const TS*)b;
  a.G = (const TS*)G; a.h = (const TS*)h; a.pen = (const TS*)pen;
  a.x = (TS*)out[0]; a.t = (TS*)out[1]; a.s = (TS*)out[2]; a.s2 = (TS*)out[3];
  a.z = (TS*)out[4]; a.z2 = (TS*)out[5]; a.y = (TS*)out[6];
  a.conv = conv; a.iters = iters;
  if (relax) {
    a.xr = (TS*)rout[0]; a.tr = (TS*)rout[1]; a.sr = (TS*)rout[2]; a.s2r = (TS*)rout[3];
    a.zr = (TS*)rout[4]; a.z2r = (TS*)rout[5]; a.yr = (TS*)rout[6];
    a.rconv = rconv; a.riters = riters;
  }
  a.tol = tol; a.kappa = kappa; a.flags = flags;
  return dispatch_nt(n, mi, [&](auto nt) { return run_e_forward<decltype(nt)::value, TS>(a, mode, relax, (hipStream_t)stream); });
}

template <typename TS>
qpx_status e_backward_impl(int mode, int64_t B, int n, int me, int mi, const void* Q, const void* q, const void* A,
                           const void* b, const void* G, const void* h, const void* pen, const void* const* rel,
                           const void* g, int flags, void* const* grads, void* ws, void* stream) {
  auto a = make_args<TS>(B, n, me, mi, ws);
  a.Q = (const TS*)Q; a.q = (const TS*)q; a.A = (const TS*)A; a.b = (const TS*)b;
  a.G = (const TS*)G; a.h = (const TS*)h; a.pen = (const TS*)pen;
  a.xr_in = (const TS*)rel[0]; a.tr_in = (const TS*)rel[1]; a.sr_in = (const TS*)rel[2];
  a.s2r_in = (const TS*)rel[3]; a.zr_in = (const TS*)rel[4]; a.z2r_in = (const TS*)rel[5];
  a.yr_in = (const TS*)rel[6]; a.g = (const TS*)g; a.flags = flags;
  a.dQ = (TS*)grads[0]; a.dq = (TS*)grads[1]; a.dA = (TS*)grads[2]; a.db = (TS*)grads[3];
  a.dG = (TS*)grads[4]; a.dh = (TS*)grads[5];
  return dispatch_nt(n, mi, [&](auto nt) { return run_e_backward<decltype(nt)::value, TS>(a, mode, (hipStream_t)stream); });
}

bool any_null(std::initializer_list<const void*> ps) {
  for (auto p : ps) if (!p) return true;
  return false;
}

// ------------------------------ the entry point that allocates nothing -------------------------
// Scratch of qpx_solve_qp_relaxed_noalloc: a file-scope device array, one copy on every GPU the
// library runs on, held until the process ends.  Host code never passes the array's name to a
// kernel (on the host it is only a stand-in for the symbol): it uses the per-device address from
// hipGetSymbolAddress, taken once per device on the entry point's first call there.
constexpr size_t kNoAllocScratchBytes = size_t(256) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

qpx_status scratch_of_current_device(void** out) {
  int dev = -1;
  QPX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(QPX_ERR_UNSUPPORTED, "device index %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(QPX_ERR_HIP, "hipGetSymbolAddress of the scratch array failed on device %d: %s", dev,
                  hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return QPX_OK;
}

}  // namespace

extern "C" {

QPX_API int32_t qpx_version(void) { return 100; }

QPX_API size_t qpx_noalloc_scratch_bytes(void) { return kNoAllocScratchBytes; }

QPX_API qpx_status qpx_solve_qp_relaxed_noalloc(int32_t dtype, int64_t batch, int32_t n, int32_t m_eq,
                                                int32_t m_ineq, const void* Q, const void* q, const void* A,
                                                const void* b, const void* G, const void* h, double solver_tol,
                                                double target_kappa, void* x, void* s, void* z, void* y,
                                                int32_t* converged, int32_t* iters, void* xr, void* sr, void* zr,
                                                void* yr, int32_t* relax_converged, int32_t* relax_iters,
                                                void* stream) {
  qpx::host::clear_error();
  if (dtype != QPX_F32 && dtype != QPX_F64) return fail(QPX_ERR_ARG, "dtype must be QPX_F32 or QPX_F64 (got %d)", dtype);
  if (batch < 0 || n < 1 || m_eq < 0 || m_ineq < 1)
    return fail(QPX_ERR_ARG, "invalid sizes batch=%lld n=%d m_eq=%d m_ineq=%d (m_ineq >= 1)", (long long)batch, n,
                m_eq, m_ineq);
  if (batch > 0x7fffffffLL) return fail(QPX_ERR_UNSUPPORTED, "batch %lld exceeds 2^31-1", (long long)batch);
  if (batch > 0 && any_null({Q, q, G, h, x, s, z, converged, iters, xr, sr, zr, relax_converged, relax_iters}))
    return fail(QPX_ERR_ARG, "null pointer argument");
  if (batch > 0 && m_eq > 0 && any_null({A, b, y, yr})) return fail(QPX_ERR_ARG, "null equality pointer");
  // per-problem words: the wavefront kernels keep only the m x m constant matrix and the split
  // launch's hand-over state; the other kernels their whole workspace view
  qpx::host::WaveVariant wv;
  const int ts = dtype == QPX_F32 ? 4 : 8;
  const int64_t stride = qpx::host::wave_variant(n, m_eq, m_ineq, ts, &wv)
                             ? qpx::round_up(qpx::wave::split_words(n, m_eq), 64)
                             : stride_words(n, m_eq, m_ineq);
  const size_t need = kCtlBytes + (size_t)batch * (size_t)stride * sizeof(double);
  if (need > kNoAllocScratchBytes)
    return fail(QPX_ERR_UNSUPPORTED,
                "batch %lld of n=%d m_eq=%d m_ineq=%d needs %zu bytes of scratch, beyond the %zu of the no-allocation "
                "entry point (use qpx_solve_qp_relaxed with a workspace)",
                (long long)batch, n, m_eq, m_ineq, need, kNoAllocScratchBytes);
  if (batch == 0) return QPX_OK;
  void* ws = nullptr;
  qpx_status st = scratch_of_current_device(&ws);
  if (st != QPX_OK) return st;
  if (dtype == QPX_F32)
    return forward_impl<float>(QPX_FUSED, batch, n, m_eq, m_ineq, Q, q, A, b, G, h, solver_tol, target_kappa, true, x,
                               s, z, y, converged, iters, xr, sr, zr, yr, relax_converged, relax_iters, ws, stream,
                               stride);
  return forward_impl<double>(QPX_FUSED, batch, n, m_eq, m_ineq, Q, q, A, b, G, h, solver_tol, target_kappa, true, x,
                              s, z, y, converged, iters, xr, sr, zr, yr, relax_converged, relax_iters, ws, stream,
                              stride);
}

QPX_API const char* qpx_last_error(void) { return qpx::host::last_error(); }

QPX_API size_t qpx_workspace_size(int64_t batch, int32_t n, int32_t m_eq, int32_t m_ineq) {
  if (batch < 0 || n < 1 || m_eq < 0 || m_ineq < 0) return 0;
  return ws_bytes(batch, n, m_eq, m_ineq);
}

QPX_API qpx_status qpx_solve_qp(int32_t dtype, int32_t mode, int64_t batch, int32_t n, int32_t m_eq,
                                int32_t m_ineq, const void* Q, const void* q, const void* A,
                                const void* b, const void* G, const void* h, double solver_tol,
                                void* x, void* s, void* z, void* y, int32_t* converged,
                                int32_t* iters, void* workspace, size_t workspace_bytes,
                                void* stream) {
  qpx::host::clear_error();
  qpx_status st = check_dims(dtype, mode, batch, n, m_eq, m_ineq, workspace_bytes);
  if (st != QPX_OK) return st;
  if (batch > 0 && any_null({Q, q, G, h, x, s, z, converged, iters, workspace}))
    return fail(QPX_ERR_ARG, "null pointer argument");
  if (batch > 0 && m_eq > 0 && any_null({A, b, y})) return fail(QPX_ERR_ARG, "null equality pointer");
  if (dtype == QPX_F32)
    return forward_impl<float>(mode, batch, n, m_eq, m_ineq, Q, q, A, b, G, h, solver_tol, 0.0, false, x, s, z, y,
                               converged, iters, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, workspace, stream);
  return forward_impl<double>(mode, batch, n, m_eq, m_ineq, Q, q, A, b, G, h, solver_tol, 0.0, false, x, s, z, y,
                              converged, iters, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, workspace, stream);
}

QPX_API qpx_status qpx_solve_qp_relaxed(int32_t dtype, int32_t mode, int64_t batch, int32_t n,
                                        int32_t m_eq, int32_t m_ineq, const void* Q, const void* q,
                                        const void* A, const void* b, const void* G, const void* h,
                                        double solver_tol, double target_kappa, void* x, void* s,
                                        void* z, void* y, int32_t* converged, int32_t* iters,
                                        void* xr, void* sr, void* zr, void* yr,
                                        int32_t* relax_converged, int32_t* relax_iters,
                                        void* workspace, size_t workspace_bytes, void* stream) {
  qpx::host::clear_error();
  qpx_status st = check_dims(dtype, mode, batch, n, m_eq, m_ineq, workspace_bytes);
  if (st != QPX_OK) return st;
  if (batch > 0 && any_null({Q, q, G, h, x, s, z, converged, iters, xr, sr, zr, relax_converged, relax_iters, workspace}))
    return fail(QPX_ERR_ARG, "null pointer argument");
  if (batch > 0 && m_eq > 0 && any_null({A, b, y, yr})) return fail(QPX_ERR_ARG, "null equality pointer");
  if (dtype == QPX_F32)
    return forward_impl<float>(mode, batch, n, m_eq, m_ineq, Q, q, A, b, G, h, solver_tol, target_kappa, true, x, s, z,
                               y, converged, iters, xr, sr, zr, yr, relax_converged, relax_iters, workspace, stream);
  return forward_impl<double>(mode, batch, n, m_eq, m_ineq, Q, q, A, b, G, h, solver_tol, target_kappa, true, x, s, z,
                              y, converged, iters, xr, sr, zr, yr, relax_converged, relax_iters, workspace, stream);
}

QPX_API qpx_status qpx_diff_qp(int32_t dtype, int32_t mode, int64_t batch, int32_t n, int32_t m_eq,
                               int32_t m_ineq, const void* Q, const void* q, const void* A,
                               const void* b, const void* G, const void* h, const void* xr,
                               const void* sr, const void* zr, const void* yr, const void* dl_dx,
                               int32_t flags, void* dQ, void* dq, void* dA, void* db, void* dG, void* dh,
                               void* workspace, size_t workspace_bytes, void* stream) {
  qpx::host::clear_error();
  qpx_status st = check_dims(dtype, mode, batch, n, m_eq, m_ineq, workspace_bytes);
  if (st != QPX_OK) return st;
  if (batch > 0 && any_null({Q, q, G, h, xr, sr, zr, dl_dx, dQ, dq, dG, dh, workspace}))
    return fail(QPX_ERR_ARG, "null pointer argument");
  if (batch > 0 && m_eq > 0 && any_null({A, b, yr, dA, db})) return fail(QPX_ERR_ARG, "null equality pointer");
  if (dtype == QPX_F32)
    return backward_impl<float>(mode, batch, n, m_eq, m_ineq, Q, q, A, b, G, h, xr, sr, zr, yr, dl_dx, flags, dQ, dq, dA,
                                db, dG, dh, workspace, stream);
  return backward_impl<double>(mode, batch, n, m_eq, m_ineq, Q, q, A, b, G, h, xr, sr, zr, yr, dl_dx, flags, dQ, dq, dA,
                               db, dG, dh, workspace, stream);
}

QPX_API qpx_status qpx_solve_qp_special(int32_t dtype, int64_t batch, int32_t n, int32_t m_eq, const void* Q,
                                        const void* q, const void* A, const void* b, void* x, void* y,
                                        int32_t* converged, int32_t* iters, void* workspace,
                                        size_t workspace_bytes, void* stream) {
  qpx::host::clear_error();
  if (dtype != QPX_F32 && dtype != QPX_F64) return fail(QPX_ERR_ARG, "dtype must be QPX_F32 or QPX_F64 (got %d)", dtype);
  if (batch < 0 || n < 1 || m_eq < 0) return fail(QPX_ERR_ARG, "invalid sizes batch=%lld n=%d m_eq=%d", (long long)batch, n, m_eq);
  if (batch > 0x7fffffffLL) return fail(QPX_ERR_UNSUPPORTED, "batch %lld exceeds 2^31-1", (long long)batch);
  // the G region of the workspace view holds 2 (m_eq + 1) columns of length n: sized by m_ineq
  const int mi = 2 * (m_eq + 1);
  const size_t need = ws_bytes(batch, n, m_eq, mi);
  if (workspace_bytes < need) return fail(QPX_ERR_WORKSPACE, "workspace of %zu bytes is smaller than the %zu required", workspace_bytes, need);
  if (batch > 0 && any_null({Q, q, x, converged, iters, workspace})) return fail(QPX_ERR_ARG, "null pointer argument");
  if (batch > 0 && m_eq > 0 && any_null({A, b, y})) return fail(QPX_ERR_ARG, "null equality pointer");
  if (batch == 0) return QPX_OK;
  auto launch = [&](auto tag) -> qpx_status {
    using TS = decltype(tag);
    auto a = make_args<TS>(batch, n, m_eq, mi, workspace);
    a.Q = (const TS*)Q; a.q = (const TS*)q; a.A = (const TS*)A; a.b = (const TS*)b;
    a.x = (TS*)x; a.y = (TS*)y; a.conv = converged; a.iters = iters;
    return dispatch_nt(n, 1, [&](auto nt) {
      hipLaunchKernelGGL((qpx::k_special<decltype(nt)::value, TS>), dim3((unsigned)batch), dim3(decltype(nt)::value), 0,
                         (hipStream_t)stream, a);
      QPX_HIP(hipGetLastError());
      return QPX_OK;
    });
  };
  return dtype == QPX_F32 ? launch(float()) : launch(double());
}

QPX_API size_t qpx_special_workspace_size(int64_t batch, int32_t n, int32_t m_eq) {
  if (batch < 0 || n < 1 || m_eq < 0) return 0;
  return ws_bytes(batch, n, m_eq, 2 * (m_eq + 1));
}

QPX_API qpx_status qpx_solve_qp_elastic(int32_t dtype, int32_t mode, int64_t batch, int32_t n, int32_t m_eq,
                                        int32_t m_ineq, const void* Q, const void* q, const void* A,
                                        const void* b, const void* G, const void* h, const void* penalty,
                                        double solver_tol, void* x, void* t, void* s1, void* s2, void* z1,
                                        void* z2, void* y, int32_t* converged, int32_t* iters, void* workspace,
                                        size_t workspace_bytes, void* stream) {
  qpx::host::clear_error();
  qpx_status st = check_dims(dtype, mode, batch, n, m_eq, m_ineq, workspace_bytes);
  if (st != QPX_OK) return st;
  if (batch > 0 && any_null({Q, q, G, h, penalty, x, t, s1, s2, z1, z2, converged, iters, workspace}))
    return fail(QPX_ERR_ARG, "null pointer argument");
  if (batch > 0 && m_eq > 0 && any_null({A, b, y})) return fail(QPX_ERR_ARG, "null equality pointer");
  void* out[7] = {x, t, s1, s2, z1, z2, y};
  if (dtype == QPX_F32)
    return e_forward_impl<float>(mode, batch, n, m_eq, m_ineq, Q, q, A, b, G, h, penalty, solver_tol, 0.0, 0, false,
                                 out, converged, iters, nullptr, nullptr, nullptr, workspace, stream);
  return e_forward_impl<double>(mode, batch, n, m_eq, m_ineq, Q, q, A, b, G, h, penalty, solver_tol, 0.0, 0, false,
                                out, converged, iters, nullptr, nullptr, nullptr, workspace, stream);
}

QPX_API qpx_status qpx_solve_qp_elastic_relaxed(int32_t dtype, int32_t mode, int64_t batch, int32_t n, int32_t m_eq,
                                                int32_t m_ineq, const void* Q, const void* q, const void* A,
                                                const void* b, const void* G, const void* h, const void* penalty,
                                                double solver_tol, double target_kappa, int32_t flags, void* x,
                                                void* t, void* s1, void* s2, void* z1, void* z2, void* y,
                                                int32_t* converged, int32_t* iters, void* xr, void* tr, void* s1r,
                                                void* s2r, void* z1r, void* z2r, void* yr, int32_t* relax_converged,
                                                int32_t* relax_iters, void* workspace, size_t workspace_bytes,
                                                void* stream) {
  qpx::host::clear_error();
  qpx_status st = check_dims(dtype, mode, batch, n, m_eq, m_ineq, workspace_bytes);
  if (st != QPX_OK) return st;
  if (batch > 0 && any_null({Q, q, G, h, penalty, x, t, s1, s2, z1, z2, converged, iters, xr, tr, s1r, s2r, z1r, z2r,
                             relax_converged, relax_iters, workspace}))
    return fail(QPX_ERR_ARG, "null pointer argument");
  if (batch > 0 && m_eq > 0 && any_null({A, b, y, yr})) return fail(QPX_ERR_ARG, "null equality pointer");
  void* out[7] = {x, t, s1, s2, z1, z2, y};
  void* rout[7] = {xr, tr, s1r, s2r, z1r, z2r, yr};
  if (dtype == QPX_F32)
    return e_forward_impl<float>(mode, batch, n, m_eq, m_ineq, Q, q, A, b, G, h, penalty, solver_tol, target_kappa,
                                 flags, true, out, converged, iters, rout, relax_converged, relax_iters, workspace,
                                 stream);
  return e_forward_impl<double>(mode, batch, n, m_eq, m_ineq, Q, q, A, b, G, h, penalty, solver_tol, target_kappa,
                                flags, true, out, converged, iters, rout, relax_converged, relax_iters, workspace,
                                stream);
}

QPX_API qpx_status qpx_diff_qp_elastic(int32_t dtype, int32_t mode, int64_t batch, int32_t n, int32_t m_eq,
                                       int32_t m_ineq, const void* Q, const void* q, const void* A, const void* b,
                                       const void* G, const void* h, const void* penalty, const void* xr,
                                       const void* tr, const void* s1r, const void* s2r, const void* z1r,
                                       const void* z2r, const void* yr, const void* dl_dx, int32_t flags, void* dQ,
                                       void* dq, void* dA, void* db, void* dG, void* dh, void* workspace,
                                       size_t workspace_bytes, void* stream) {
  qpx::host::clear_error();
  qpx_status st = check_dims(dtype, mode, batch, n, m_eq, m_ineq, workspace_bytes);
  if (st != QPX_OK) return st;
  if (batch > 0 && any_null({Q, q, G, h, penalty, xr, tr, s1r, s2r, z1r, z2r, dl_dx, dQ, dq, dG, dh, workspace}))
    return fail(QPX_ERR_ARG, "null pointer argument");
  if (batch > 0 && m_eq > 0 && any_null({A, b, yr, dA, db})) return fail(QPX_ERR_ARG, "null equality pointer");
  const void* rel[7] = {xr, tr, s1r, s2r, z1r, z2r, yr};
  void* grads[6] = {dQ, dq, dA, db, dG, dh};
  if (dtype == QPX_F32)
    return e_backward_impl<float>(mode, batch, n, m_eq, m_ineq, Q, q, A, b, G, h, penalty, rel, dl_dx, flags, grads,
                                  workspace, stream);
  return e_backward_impl<double>(mode, batch, n, m_eq, m_ineq, Q, q, A, b, G, h, penalty, rel, dl_dx, flags, grads,
                                 workspace, stream);
}

QPX_API qpx_status qpx_initialize(int32_t dtype, int64_t batch, int32_t n, int32_t m_eq, int32_t m_ineq,
                                  const void* Q, const void* q, const void* A, const void* b, const void* G,
                                  const void* h, void* x, void* s, void* z, void* y, void* workspace,
                                  size_t workspace_bytes, void* stream) {
  qpx::host::clear_error();
  qpx_status st = check_dims(dtype, QPX_FUSED, batch, n, m_eq, m_ineq, workspace_bytes);
  if (st != QPX_OK) return st;
  if (batch > 0 && any_null({Q, q, G, h, x, s, z, workspace})) return fail(QPX_ERR_ARG, "null pointer argument");
  if (batch > 0 && m_eq > 0 && any_null({A, b, y})) return fail(QPX_ERR_ARG, "null equality pointer");
  if (batch == 0) return QPX_OK;
  auto launch = [&](auto tag) -> qpx_status {
    using TS = decltype(tag);
    auto a = make_args<TS>(batch, n, m_eq, m_ineq, workspace);
    a.Q = (const TS*)Q; a.q = (const TS*)q; a.A = (const TS*)A; a.b = (const TS*)b;
    a.G = (const TS*)G; a.h = (const TS*)h;
    a.x = (TS*)x; a.s = (TS*)s; a.z = (TS*)z; a.y = (TS*)y;
    return dispatch_nt(n, m_ineq, [&](auto nt) {
      const int lds = qpx::generic_lds_bytes(n, m_eq, decltype(nt)::value);
      hipLaunchKernelGGL((qpx::k_initialize<decltype(nt)::value, TS>), dim3((unsigned)batch), dim3(decltype(nt)::value),
                         lds, (hipStream_t)stream, a);
      QPX_HIP(hipGetLastError());
      return QPX_OK;
    });
  };
  return dtype == QPX_F32 ? launch(float()) : launch(double());
}

QPX_API qpx_status qpx_initialize_elastic(int32_t dtype, int64_t batch, int32_t n, int32_t m_eq, int32_t m_ineq,
                                          const void* Q, const void* q, const void* A, const void* b, const void* G,
                                          const void* h, const void* penalty, void* x, void* t, void* s1, void* s2,
                                          void* z1, void* z2, void* y, void* workspace, size_t workspace_bytes,
                                          void* stream) {
  qpx::host::clear_error();
  qpx_status st = check_dims(dtype, QPX_FUSED, batch, n, m_eq, m_ineq, workspace_bytes);
  if (st != QPX_OK) return st;
  if (batch > 0 && any_null({Q, q, G, h, penalty, x, t, s1, s2, z1, z2, workspace}))
    return fail(QPX_ERR_ARG, "null pointer argument");
  if (batch > 0 && m_eq > 0 && any_null({A, b, y})) return fail(QPX_ERR_ARG, "null equality pointer");
  if (batch == 0) return QPX_OK;
  auto launch = [&](auto tag) -> qpx_status {
    using TS = decltype(tag);
    auto a = make_args<TS>(batch, n, m_eq, m_ineq, workspace);
    a.Q = (const TS*)Q; a.q = (const TS*)q; a.A = (const TS*)A; a.b = (const TS*)b;
    a.G = (const TS*)G; a.h = (const TS*)h; a.pen = (const TS*)penalty;
    a.x = (TS*)x; a.t = (TS*)t; a.s = (TS*)s1; a.s2 = (TS*)s2; a.z = (TS*)z1; a.z2 = (TS*)z2; a.y = (TS*)y;
    return dispatch_nt(n, m_ineq, [&](auto nt) {
      const int lds = qpx::generic_lds_bytes(n, m_eq, decltype(nt)::value);
      hipLaunchKernelGGL((qpx::k_e_initialize<decltype(nt)::value, TS>), dim3((unsigned)batch),
                         dim3(decltype(nt)::value), lds, (hipStream_t)stream, a);
      QPX_HIP(hipGetLastError());
      return QPX_OK;
    });
  };
  return dtype == QPX_F32 ? launch(float()) : launch(double());
}

QPX_API qpx_status qpx_solve_qp_trace(int32_t dtype, int64_t batch, int32_t n, int32_t m_eq, int32_t m_ineq,
                                      const void* Q, const void* q, const void* A, const void* b, const void* G,
                                      const void* h, double solver_tol, int32_t flags, void* x, void* s, void* z,
                                      void* y, int32_t* converged, int32_t* iters, double* trace, void* workspace,
                                      size_t workspace_bytes, void* stream) {
  qpx::host::clear_error();
  qpx_status st = check_dims(dtype, QPX_FUSED, batch, n, m_eq, m_ineq, workspace_bytes);
  if (st != QPX_OK) return st;
  if (batch > 0 && any_null({Q, q, G, h, x, s, z, converged, iters, trace, workspace}))
    return fail(QPX_ERR_ARG, "null pointer argument");
  if (batch > 0 && m_eq > 0 && any_null({A, b, y})) return fail(QPX_ERR_ARG, "null equality pointer");
  if (flags & ~QPX_TRACE_DEBUG_STEP) return fail(QPX_ERR_ARG, "unknown flags 0x%x", flags);
  if (batch == 0) return QPX_OK;
  auto launch = [&](auto tag) -> qpx_status {
    using TS = decltype(tag);
    auto a = make_args<TS>(batch, n, m_eq, m_ineq, workspace);
    a.Q = (const TS*)Q; a.q = (const TS*)q; a.A = (const TS*)A; a.b = (const TS*)b;
    a.G = (const TS*)G; a.h = (const TS*)h;
    a.x = (TS*)x; a.s = (TS*)s; a.z = (TS*)z; a.y = (TS*)y; a.conv = converged; a.iters = iters;
    a.tol = solver_tol; a.flags = flags; a.trace = trace;
    return dispatch_nt(n, m_ineq, [&](auto nt) {
      const int lds = qpx::generic_lds_bytes(n, m_eq, decltype(nt)::value);
      hipLaunchKernelGGL((qpx::k_fused_solve<decltype(nt)::value, TS>), dim3((unsigned)batch),
                         dim3(decltype(nt)::value), lds, (hipStream_t)stream, a, 0);
      QPX_HIP(hipGetLastError());
      return QPX_OK;
    });
  };
  return dtype == QPX_F32 ? launch(float()) : launch(double());
}

}  // extern "C"
