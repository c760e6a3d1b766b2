// Device building blocks of the batched PDIP solver: one thread block owns one QP.
//
// Every function below is a block-cooperative stage of qpax's algorithm, written for a block of
// NT threads (NT a multiple of 64).  All IPM arithmetic is float64 whatever the storage type of
// the caller's arrays: the reduced KKT matrix H = Q + G^T diag(z/s) G reaches condition numbers
// far beyond 1/eps(float32) as the iterates approach the boundary, and on gfx950 (packed FP32
// disabled) a float64 FMA issues at the rate of a float32 one.
//
// Linear algebra replaces the reference's QR (qpax/pdip.py:10-12,38-46) by Cholesky factors of
// the same symmetric positive definite matrices; the reduction itself (cvxopt, Algorithm 1,
// qpax/pdip.py:49-63) is unchanged.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace qpx {

constexpr int kMaxIters = 30;          // qpax/pdip.py:232, pdip_relaxed.py:76
constexpr double kPivotRel = 1e-14;    // pivot floor relative to the matrix's own diagonal entry
constexpr double kHugePivot = 1e64;    // a pivot below the floor is replaced by this (drops the direction)

// jnp.min / jnp.max propagate NaN; so do these.
__device__ __forceinline__ double nmin(double a, double b) { return (a < b || a != a) ? a : b; }
__device__ __forceinline__ double nmax(double a, double b) { return (a > b || a != a) ? a : b; }

// Copies of a lane index or of a uniform size that the compiler cannot hoist out of the
// iteration loops: lane masks and addresses derived from them are recomputed per stage instead of
// being kept live in registers across the whole solve.
// (QPX_PIN_V / QPX_PIN_S and QPX_DYN_LDS are predefined only by the host build of these sources
// that the CPU bounds tests run.)
#ifndef QPX_PIN_V
#define QPX_PIN_V(x) asm volatile("" : "+v"(x))
#define QPX_PIN_S(x) asm volatile("" : "+s"(x))
#endif
// A pointer's global-memory view: loads through it are `global_load`s, counted by vmcnt alone (a
// generic pointer's `flat_load`s also count in lgkmcnt, so every LDS wait would wait for them too).
#ifndef QPX_GLOBAL_AS
#define QPX_GLOBAL_AS __attribute__((address_space(1)))
#endif
#ifndef QPX_LDS_AS
#define QPX_LDS_AS __attribute__((address_space(3)))
#endif
template <typename T>
__device__ __forceinline__ const QPX_GLOBAL_AS T* as_global(const T* p) {
  return (const QPX_GLOBAL_AS T*)p;
}
// The same for stores (global stores count in vmcnt only; flat stores also hold up LDS waits).
template <typename T>
__device__ __forceinline__ QPX_GLOBAL_AS T* as_global_w(T* p) {
  return (QPX_GLOBAL_AS T*)p;
}
// A pointer that every lane holds alike (a workgroup's matrix), moved to scalar registers: loads
// through it then take a scalar base and a 32-bit per-lane offset, one register per address
// instead of a 64-bit sum.
template <typename T>
__device__ __forceinline__ const QPX_GLOBAL_AS T* uniform_global(const T* p) {
  const uint64_t u = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
  return (const QPX_GLOBAL_AS T*)(((uint64_t)hi << 32) | lo);
}
template <typename T>
__device__ __forceinline__ QPX_GLOBAL_AS T* uniform_global_w(T* p) {
  const uint64_t u = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
  return (QPX_GLOBAL_AS T*)(((uint64_t)hi << 32) | lo);
}
// A generic pointer parameter known to address LDS: with this assumption the compiler issues LDS
// instructions through it instead of flat ones, which count against both the LDS and the
// vector-memory wait counters.
#ifndef QPX_ASSUME_LDS
#if defined(__HIP_DEVICE_COMPILE__)
#define QPX_ASSUME_LDS(p) __builtin_assume(__builtin_amdgcn_is_shared((const void*)(p)))
#else
#define QPX_ASSUME_LDS(p) ((void)0)  // the host pass of the same source
#endif
#endif
// Lane `lane`'s value of v (a uniform result).
__device__ __forceinline__ double readlane_d(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}
// The workgroup's dynamic LDS as `unsigned char* name`.
#ifndef QPX_DYN_LDS
#define QPX_DYN_LDS(name) extern __shared__ __align__(16) unsigned char name[]
#endif
__device__ __forceinline__ int opaque_v(int x) {
  QPX_PIN_V(x);
  return x;
}
__device__ __forceinline__ int opaque_s(int x) {
  x = __builtin_amdgcn_readfirstlane(x);
  QPX_PIN_S(x);
  return x;
}

struct OpMin { __device__ double operator()(double a, double b) const { return nmin(a, b); } };
struct OpMax { __device__ double operator()(double a, double b) const { return nmax(a, b); } };
struct OpSum { __device__ double operator()(double a, double b) const { return a + b; } };

// Reduction over the block; every thread receives the result.  `red` is LDS scratch of NT/64
// doubles.  Contains barriers: call from uniform control flow only.
template <int NT, class Op>
__device__ __forceinline__ double block_reduce(double v, double* red, Op op) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o, 64));
  if constexpr (NT > 64) {
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    v = red[0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) v = op(v, red[w]);
  }
  return v;
}

// --------------------------------------------------------------------------------------------
// Per-problem view of the float64 workspace.
// --------------------------------------------------------------------------------------------
enum Ctl : int { kConv = 0, kIter = 1, kActive = 2, kRelaxConv = 3, kRelaxIter = 4, kCtlWords = 8 };

struct View {
  int n, me, mi;
  // sanitised problem: Q symmetrised, rows of G and entries of h zeroed where h is +-inf
  double *Q, *G, *A, *q, *b, *h, *hinf;
  // iterate
  double *x, *s, *z, *y;
  // factorisation: H (n x n, lower = L), W = L^{-1} A^T (n x me, row-major), F (me x me, lower)
  double *H, *W, *F, *hdiag, *hrcp, *fdiag, *frcp;
  // work vectors
  double *D, *dx, *ds, *dz, *dy, *dsa, *dza, *r1, *r2, *r3, *r4, *u, *t;
  double *scal;  // [0]=sigma*mu, [1]=alpha, [2]=kappa
  int* ctl;
};

__host__ __device__ inline int64_t round_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

// Number of float64 words of workspace one problem of size (n, me, mi) occupies.
__host__ __device__ inline int64_t view_words(int n, int me, int mi) {
  int64_t w = 0;
  auto add = [&](int64_t k) { w += round_up(k > 0 ? k : 1, 2); };
  add((int64_t)n * n); add((int64_t)mi * n); add((int64_t)me * n); add(n); add(me); add(mi); add(mi);
  add(n); add(mi); add(mi); add(me);
  add((int64_t)n * n); add((int64_t)n * me); add((int64_t)me * me); add(n); add(n); add(me); add(me);
  add(mi); add(n); add(mi); add(mi); add(me); add(mi); add(mi); add(n); add(mi); add(mi); add(me); add(n); add(me);
  add(8);
  add(kCtlWords / 2);
  return round_up(w, 64);
}

__host__ __device__ inline View carve(double* base, int n, int me, int mi) {
  View v;
  v.n = n; v.me = me; v.mi = mi;
  double* p = base;
  auto take = [&](int64_t k) { double* r = p; p += round_up(k > 0 ? k : 1, 2); return r; };
  v.Q = take((int64_t)n * n); v.G = take((int64_t)mi * n); v.A = take((int64_t)me * n);
  v.q = take(n); v.b = take(me); v.h = take(mi); v.hinf = take(mi);
  v.x = take(n); v.s = take(mi); v.z = take(mi); v.y = take(me);
  v.H = take((int64_t)n * n); v.W = take((int64_t)n * me); v.F = take((int64_t)me * me);
  v.hdiag = take(n); v.hrcp = take(n); v.fdiag = take(me); v.frcp = take(me);
  v.D = take(mi); v.dx = take(n); v.ds = take(mi); v.dz = take(mi); v.dy = take(me);
  v.dsa = take(mi); v.dza = take(mi); v.r1 = take(n); v.r2 = take(mi); v.r3 = take(mi); v.r4 = take(me);
  v.u = take(n); v.t = take(me);
  v.scal = take(8);
  v.ctl = reinterpret_cast<int*>(take(kCtlWords / 2));
  return v;
}

// --------------------------------------------------------------------------------------------
// Stage: load and sanitise the problem (qpax/pdip.py:166-178 inf rows, :223 symmetrisation).
// --------------------------------------------------------------------------------------------
template <int NT, typename TS>
__device__ void stage_load(const View& v, const TS* Q, const TS* q, const TS* A, const TS* b,
                           const TS* G, const TS* h) {
  const int n = v.n, me = v.me, mi = v.mi, tid = threadIdx.x;
  for (int e = tid; e < n * n; e += NT) {
    const int i = e / n, j = e - i * n;
    v.Q[e] = 0.5 * ((double)Q[e] + (double)Q[j * n + i]);
  }
  for (int k = tid; k < mi; k += NT) {
    const double hk = (double)h[k];
    const bool inf = isinf(hk);
    v.h[k] = inf ? 0.0 : hk;
    v.hinf[k] = inf ? 1.0 : 0.0;
  }
  for (int i = tid; i < n; i += NT) v.q[i] = (double)q[i];
  for (int c = tid; c < me; c += NT) v.b[c] = (double)b[c];
  for (int e = tid; e < me * n; e += NT) v.A[e] = (double)A[e];
  for (int e = tid; e < mi * n; e += NT) {
    const int k = e / n;
    // diag(1 * ~mask) @ G: a masked row becomes 0 * G (pdip.py:176)
    v.G[e] = isinf((double)h[k]) ? 0.0 * (double)G[e] : (double)G[e];
  }
  if (tid < kCtlWords) v.ctl[tid] = 0;
  __syncthreads();
}

// --------------------------------------------------------------------------------------------
// Cholesky of a symmetric matrix held in the lower triangle of M (m x m, row-major, leading
// dimension m).  Overwrites the lower triangle with L.  diag0/rcp receive the original diagonal
// and 1/L_jj.  A pivot that is not above kPivotRel times its original diagonal entry (or is
// NaN) is replaced by kHugePivot, which zeroes that component of later solves.
// --------------------------------------------------------------------------------------------
template <int NT>
__device__ void block_cholesky(double* M, int m, double* diag0, double* rcp) {
  const int tid = threadIdx.x;
  for (int i = tid; i < m; i += NT) diag0[i] = M[i * m + i];
  __syncthreads();
  for (int j = 0; j < m; ++j) {
    const double d = M[j * m + j];
    const bool ok = d > kPivotRel * fabs(diag0[j]);
    const double ljj = ok ? sqrt(d) : kHugePivot;
    const double inv = 1.0 / ljj;
    for (int i = j + 1 + tid; i < m; i += NT) M[i * m + j] *= inv;
    __syncthreads();
    if (tid == 0) { M[j * m + j] = ljj; rcp[j] = inv; }
    const int r = m - j - 1;  // trailing block is r x r, rows/cols j+1..m-1
    for (int e = tid; e < r * r; e += NT) {
      const int a = e / r, c = e - a * r;
      if (c <= a) {
        const int i = j + 1 + a, k = j + 1 + c;
        M[i * m + k] -= M[i * m + j] * M[k * m + j];
      }
    }
    __syncthreads();
  }
}

// Forward substitution L y = b for one right-hand side (m < 64: the sizes below the blocked
// sweeps); b is overwritten with y.  Lane i of wavefront 0 holds b_i and applies column j's update
// b_i -= L_ij y_j (i > j) in column order, as a column sweep over the block would, with eight
// columns' entries of L and 1/L_jj loaded ahead (clamped indices).
template <int NT>
__device__ void block_fsub_scalar(const double* L, const double* rcp, int m, double* b) {
  __syncthreads();
  if (threadIdx.x < 64) {
    const int l = threadIdx.x, lr = l < m ? l : m - 1;
    double bi = l < m ? b[l] : 0.0;
    for (int j0 = 0; j0 < m; j0 += 8) {
      double lv[8], rc[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int j = j0 + t < m ? j0 + t : m - 1;
        lv[t] = L[lr * m + j];
        rc[t] = rcp[j];
      }
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int j = j0 + t;
        if (j < m) {
          const double yj = readlane_d(bi, j) * rc[t];
          bi = (l == j) ? yj : (l > j ? bi - lv[t] * yj : bi);
        }
      }
    }
    if (l < m) b[l] = bi;
  }
  __syncthreads();
}

// Back substitution L^T x = b (m < 64); b is overwritten with x.  Lane i of wavefront 0 holds b_i;
// columns from the last, b_i -= L_ji x_j for i < j.
template <int NT>
__device__ void block_bsub_t_scalar(const double* L, const double* rcp, int m, double* b) {
  __syncthreads();
  if (threadIdx.x < 64) {
    const int l = threadIdx.x, lc = l < m ? l : m - 1;
    double bi = l < m ? b[l] : 0.0;
    for (int j0 = m - 1; j0 >= 0; j0 -= 8) {
      double lv[8], rc[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int j = j0 - t >= 0 ? j0 - t : 0;
        lv[t] = L[j * m + lc];
        rc[t] = rcp[j];
      }
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int j = j0 - t;
        if (j >= 0) {
          const double xj = readlane_d(bi, j) * rc[t];
          bi = (l == j) ? xj : (l < j ? bi - lv[t] * xj : bi);
        }
      }
    }
    if (l < m) b[l] = bi;
  }
  __syncthreads();
}

// --------------------------------------------------------------------------------------------
// Stage: factorisation (qpax/pdip.py:38-46).  D = z/s; H = Q + G^T D G; H = L L^T;
// W = L^{-1} A^T; F = W^T W = A H^{-1} A^T; F = Lf Lf^T.
// The scaling is D = z/s (kDzs), D = 1 (kDunit, the initialisation, pdip.py:17) or already in
// v.D (kDgiven, used by the elastic reduction).
// --------------------------------------------------------------------------------------------
enum DMode : int { kDzs = 0, kDunit = 1, kDgiven = 2 };

// Matrix-core G^T D G for n >= kMfmaMinN: the lower triangle of H = Q + G^T D G as 16x16 tiles
// accumulated with v_mfma_f64_16x16x4f64.  Wavefront w takes tiles w, w + W, ... (row-major over
// the lower triangle) in batches of kGtdgBatch accumulators; each batch streams G through the
// dynamic LDS in chunks of kGtdgRows rows (row stride 16 * (tiles | 1), zero-padded columns; the
// odd tile count keeps an operand load's four rows in disjoint banks) with D beside them.
constexpr int kMfmaMinN = 64;
constexpr int kMfmaMaxN = 272;
constexpr int kGtdgRows = 16;
constexpr int kGtdgBatch = 17;
typedef double d4v __attribute__((ext_vector_type(4)));

__host__ __device__ inline int gtdg_stride(int n) { return 16 * (((n + 15) / 16) | 1); }
constexpr int kPanelLd = 17;  // row stride of the Cholesky panel in LDS
// One Cholesky panel in LDS: a 16-column block column of an n x n matrix (row stride kPanelLd).
__host__ __device__ inline int panel_words(int n) { return kPanelLd * 16 * ((n + 15) / 16); }
// Dynamic LDS bytes the generic kernels need for a problem with n variables: the G chunk of the
// matrix-core G^T D G, and (aliasing it) the two panels of the blocked Cholesky.
__host__ __device__ inline int gtdg_lds_bytes(int n) {
  if (n < kMfmaMinN || n > kMfmaMaxN) return 0;
  const int a = kGtdgRows * gtdg_stride(n) + kGtdgRows, b = 2 * panel_words(n);
  return 8 * (a > b ? a : b);
}
// W = L^{-1} A^T (n x me float64) is formed in LDS when the factorisation is the tiled one, W fits
// this budget and the workgroup has 256 threads (in the 64/128-thread kernels, which also serve
// n < 64, the LDS path's registers would lower the occupancy of every launch).  It is formed inside
// the Cholesky, beside its two panels, when all three fit kFusedLdsMax; otherwise after it,
// aliasing them.
constexpr int kWLdsMax = 128 * 1024;
constexpr int kFusedLdsMax = 144 * 1024;
__host__ __device__ inline bool w_in_lds(int n, int me, int nt) {
  return nt >= 256 && n >= kMfmaMinN && n <= kMfmaMaxN && me > 0 && 8 * n * me <= kWLdsMax;
}
__host__ __device__ inline bool w_fused(int n, int me, int nt) {
  return w_in_lds(n, me, nt) && 8 * (2 * panel_words(n) + n * me) <= kFusedLdsMax;
}
// Dynamic LDS bytes of the nt-thread generic kernels for a problem of n variables and me equality
// rows.
__host__ __device__ inline int generic_lds_bytes(int n, int me, int nt) {
  int b = gtdg_lds_bytes(n);
  // the column sums' partials (colsum_lds_bytes) also live there
  const int cs = 8 * nt * (nt >= 256 ? (kMfmaMaxN + 63) / 64 : 2);
  if (n >= kMfmaMinN && n <= kMfmaMaxN && b < cs) b = cs;
  if (w_in_lds(n, me, nt) && 8 * n * me > b) b = 8 * n * me;
  if (n >= kMfmaMinN && n <= kMfmaMaxN) {
    // the blocked Cholesky: two panels, W beside them when fused, then 1/L_jj, a column buffer, the
    // reference diagonal and L_kk^{-1}
    const int c = 8 * (2 * panel_words(n) + (w_fused(n, me, nt) ? n * me : 0) + 2 * n + 16 + 16 * kPanelLd);
    if (c > b) b = c;
  }
  return b;
}

// Sum over each group of 16 lanes (DPP row butterflies); every lane of the group receives it.
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double sum16(double v) {
  v += dpp_mov<0xB1>(v);
  v += dpp_mov<0x4E>(v);
  v += dpp_mov<0x141>(v);
  v += dpp_mov<0x140>(v);
  return v;
}

// out[k] = sum_j M[k n + j] x[j] for every row k < rows: 16 lanes per row read it coalesced and
// meet in a DPP reduction (the large-size form of the row products, n >= kMfmaMinN).  Up to
// kMfmaMaxN columns a lane keeps its slice of x (columns cq, cq + 16, ...) in registers and each
// wavefront has two groups of four rows' loads in flight (clamped indices); wider rows loop.
template <int NT, typename MT>
__device__ __forceinline__ void rows_dot16(const MT* M, int rows, int n, const double* x, double* out) {
  constexpr int W = NT / 64;
  constexpr int S = (kMfmaMaxN + 15) / 16;
  const int tid = threadIdx.x, w = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63, rq = l >> 4, cq = l & 15;
  const QPX_GLOBAL_AS MT* Mg = as_global(M);
  if (n <= kMfmaMaxN) {
    double xs[S];
#pragma unroll
    for (int u = 0; u < S; ++u) {
      const int j = cq + 16 * u;
      xs[u] = j < n ? x[j] : 0.0;
    }
    // NG groups of four rows per wavefront in flight: eight for 4-byte elements, two for 8-byte
    constexpr int NG = sizeof(MT) <= 4 ? 8 : 2;
    for (int kb = 4 * w; kb < rows; kb += 4 * NG * W) {
      MT mv[NG][S];
#pragma unroll
      for (int g = 0; g < NG; ++g) {
        const int k0 = kb + 4 * W * g + rq, k = k0 < rows ? k0 : rows - 1;
#pragma unroll
        for (int u = 0; u < S; ++u) {
          const int j = cq + 16 * u;
          mv[g][u] = Mg[(int64_t)k * n + (j < n ? j : n - 1)];
        }
      }
#pragma unroll
      for (int g = 0; g < NG; ++g) {
        double a0 = 0.0, a1 = 0.0;
#pragma unroll
        for (int u = 0; u < S; u += 2) {
          if (cq + 16 * u < n) a0 += (double)mv[g][u] * xs[u];
          if (u + 1 < S && cq + 16 * (u + 1) < n) a1 += (double)mv[g][u + 1] * xs[u + 1];
        }
        const double v = sum16(a0 + a1);
        const int k = kb + 4 * W * g + rq;
        if (cq == 0 && k < rows) out[k] = v;
      }
    }
    return;
  }
  for (int kb = 4 * w; kb < rows; kb += 4 * W) {
    const int k = kb + rq;
    const MT* mr = M + (int64_t)(k < rows ? k : rows - 1) * n;
    double a0 = 0.0, a1 = 0.0;
    int j = cq;
    for (; j + 16 < n; j += 32) {
      a0 += (double)mr[j] * x[j];
      a1 += (double)mr[j + 16] * x[j + 16];
    }
    if (j < n) a0 += (double)mr[j] * x[j];
    const double v = sum16(a0 + a1);
    if (cq == 0 && k < rows) out[k] = v;
  }
}

// out[k] = sum_j M[k n + j] x[j] for every row k < rows, as rows_dot16 (n <= kMfmaMaxN), and from
// the same loads the column sums gcol[j] = sum_k M[k n + j] v[k]: each lane accumulates its
// columns (cq, cq + 16, ...) over its rows, and the 4 W partial sums of a column meet through
// `part` (LDS, 4 W x 16 ceil(n / 16) doubles), summed in a fixed order.  gcol may be in LDS or
// global memory.  One read of M for both products.  Contains barriers.
template <int NT, typename MT>
__device__ __forceinline__ void rows_cols16(const MT* M, int rows, int n, const double* x, double* out,
                                            const double* v, double* gcol, double* part) {
  constexpr int W = NT / 64;
  constexpr int S = (kMfmaMaxN + 15) / 16;
  const int tid = threadIdx.x, w = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63, rq = l >> 4, cq = l & 15;
  const QPX_GLOBAL_AS MT* Mg = uniform_global(M);
  const QPX_GLOBAL_AS double* vg = as_global(v);
  const int P = 16 * ((n + 15) / 16);
  double xs[S], ca[S];
#pragma unroll
  for (int u = 0; u < S; ++u) {
    const int j = cq + 16 * u;
    xs[u] = j < n ? x[j] : 0.0;
    ca[u] = 0.0;
  }
  // NG groups of four rows per wavefront in flight: four for 4-byte elements, two for 8-byte
  constexpr int NG = sizeof(MT) <= 4 ? 4 : 2;
  for (int kb = 4 * w; kb < rows; kb += 4 * NG * W) {
    MT mv[NG][S];
    double vk[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      const int k0 = kb + 4 * W * g + rq, k = k0 < rows ? k0 : rows - 1;
      const double vv = vg[(unsigned)k];
      vk[g] = k0 < rows ? vv : 0.0;
#pragma unroll
      for (int u = 0; u < S; ++u) {
        const int j = cq + 16 * u;
        mv[g][u] = Mg[(unsigned)(k * n + (j < n ? j : n - 1))];
      }
    }
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      double a0 = 0.0, a1 = 0.0;
#pragma unroll
      for (int u = 0; u < S; u += 2) {
        if (cq + 16 * u < n) a0 += (double)mv[g][u] * xs[u];
        if (u + 1 < S && cq + 16 * (u + 1) < n) a1 += (double)mv[g][u + 1] * xs[u + 1];
      }
#pragma unroll
      for (int u = 0; u < S; ++u) ca[u] += (double)mv[g][u] * vk[g];
      const double sv = sum16(a0 + a1);
      const int k = kb + 4 * W * g + rq;
      if (cq == 0 && k < rows) out[k] = sv;
    }
  }
  QPX_ASSUME_LDS(part);
  __syncthreads();
#pragma unroll
  for (int u = 0; u < S; ++u)
    if (16 * u < P) part[(4 * w + rq) * P + cq + 16 * u] = ca[u];
  __syncthreads();
  for (int c = tid; c < n; c += NT) {
    double sum = 0.0;
#pragma unroll
    for (int q = 0; q < 4 * W; ++q) sum += part[q * P + c];
    gcol[c] = sum;
  }
  __syncthreads();
}

// Column slots per lane of the column sums in the NT-thread generic kernels (n <= 96 runs on 128
// threads, n <= kMfmaMaxN on 256), and the LDS they take for the partial sums: within the
// dynamic LDS of those kernels (the G^T D G chunk, free outside the factorisation) for n >= 64.
template <int NT>
constexpr int colsum_slots() { return NT >= 256 ? (kMfmaMaxN + 63) / 64 : 2; }
template <int NT>
constexpr int colsum_lds_bytes() { return 8 * (NT / 64) * 64 * colsum_slots<NT>(); }
static_assert(colsum_lds_bytes<128>() <= 8 * (kGtdgRows * 80 + kGtdgRows), "column sums exceed the n = 64 chunk");
static_assert(colsum_lds_bytes<256>() <= 8 * (kGtdgRows * 112 + kGtdgRows), "column sums exceed the n = 97 chunk");

// Column sums acc[u] += sum_k M[k, c] v[k] for the columns c = l + 64 u (< cols <= 64 CPL) of this
// lane: the wavefronts take rows w, w + W, ..., 32 (4-byte elements) or eight rows' loads in
// flight per lane (clamped indices; rows past the end add nothing).  Partial sums per wavefront:
// finish with colsum_finish.
template <int NT, int CPL, typename T>
__device__ __forceinline__ void colsum_acc(const T* M, int rows, int cols, const double* v, double (&acc)[CPL]) {
  constexpr int W = NT / 64;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
  const QPX_GLOBAL_AS T* Mg = as_global(M);
  // RD rows' loads in flight per lane: 32 for 4-byte elements, eight for 8-byte
  constexpr int RD = sizeof(T) <= 4 ? 32 : 8;
  for (int k0 = w; k0 < rows; k0 += RD * W) {
    T m[RD][CPL];
    double vk[RD];
#pragma unroll
    for (int r = 0; r < RD; ++r) {
      const int k = k0 + r * W < rows ? k0 + r * W : rows - 1;
      vk[r] = v[k];
#pragma unroll
      for (int u = 0; u < CPL; ++u) {
        const int c = l + 64 * u;
        if (64 * u < cols) m[r][u] = Mg[(int64_t)k * cols + (c < cols ? c : cols - 1)];
      }
    }
#pragma unroll
    for (int r = 0; r < RD; ++r)
      if (k0 + r * W < rows)
#pragma unroll
        for (int u = 0; u < CPL; ++u)
          if (64 * u < cols) acc[u] += (double)m[r][u] * vk[r];
  }
}

// out[c] = init[c] + sg * (sum of the wavefronts' partial column sums), summed in wavefront order
// through `part` (LDS, NT / 64 x 64 CPL doubles); the sum of |out| entries' maximum is returned to
// every thread when want_max.  Contains barriers.
template <int NT, int CPL>
__device__ __forceinline__ void colsum_finish(const double (&acc)[CPL], int cols, const double* init, double si,
                                              double sg, double* out, double* part) {
  constexpr int W = NT / 64;
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  __syncthreads();
#pragma unroll
  for (int u = 0; u < CPL; ++u) part[w * 64 * CPL + l + 64 * u] = acc[u];
  __syncthreads();
  for (int c = tid; c < cols; c += NT) {
    double s = part[c];
#pragma unroll
    for (int ww = 1; ww < W; ++ww) s += part[ww * 64 * CPL + c];
    out[c] = si * init[c] + sg * s;
  }
  __syncthreads();
}

// G is read from `G` (the workspace copy, or for the elastic solve the caller's array, whose
// element type GT widens exactly to the copy's float64 values).
template <int NT, typename GT>
__device__ __noinline__ void gtdg_mfma(const double* Q, const double* D, double* H, int n, int mi, double* chunk,
                                       const GT* G) {
  constexpr int W = NT / 64;
  // 256-thread blocks prefetch the next chunk into registers under the MFMAs and keep 17
  // accumulators per wavefront (the 136 tiles of n = 256 in two passes); smaller blocks (n <= 96)
  // load in place and keep 8
  constexpr bool kPf = NT >= 256;
  constexpr int kBatch = kPf ? kGtdgBatch : kGtdgBatch / 2;
  constexpr int EPT = kPf ? (kGtdgRows * (16 * ((kMfmaMaxN / 16) | 1)) + NT - 1) / NT : 1;
  const int tid = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63, rq = l >> 4, cq = l & 15;
  const int nt = (n + 15) / 16, ldl = gtdg_stride(n), T = nt * (nt + 1) / 2;
  const int csz = kGtdgRows * n;  // one chunk: kGtdgRows consecutive rows of G (row stride n)
  double* cD = chunk + kGtdgRows * ldl;
  // the chunk's elements this thread moves: flat index e = tid + NT u over kGtdgRows x n, placed
  // at row e / n, column e % n of the LDS chunk (the columns n..ldl-1 are zeroed once per pass)
  int dst[EPT];
#pragma unroll
  for (int u = 0; u < EPT; ++u) {
    const int e = tid + NT * u;
    dst[u] = e < csz ? (e / n) * ldl + (e - (e / n) * n) : -1;
  }
  // the next chunk in flight: raw elements (converted when stored) loaded unconditionally through
  // clamped indices, so that every load issues before the first wait
  GT reg[EPT];
  double dreg = 0.0;
  const QPX_GLOBAL_AS GT* Gg = as_global(G);
  const int64_t last = (int64_t)mi * n - 1;
  auto fetch = [&](int k0) {
    const int64_t base = (int64_t)k0 * n;
#pragma unroll
    for (int u = 0; u < EPT; ++u) {
      const int64_t g = base + tid + NT * u;
      reg[u] = Gg[g < last ? g : last];
    }
    const int kd = k0 + (tid & (kGtdgRows - 1));
    dreg = as_global(D)[kd < mi ? kd : mi - 1];
  };
  for (int base = 0; base < T; base += W * kBatch) {
    int tI[kBatch], tJ[kBatch];
    bool ok[kBatch];
#pragma unroll
    for (int b = 0; b < kBatch; ++b) {
      const int idx = base + b * W + w;
      int I = (int)((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
      while ((I + 1) * (I + 2) / 2 <= idx) ++I;
      while (I * (I + 1) / 2 > idx) --I;
      ok[b] = idx < T;
      tI[b] = ok[b] ? I : 0;
      tJ[b] = ok[b] ? idx - I * (I + 1) / 2 : 0;
    }
    const int nb = (T - base - w + W - 1) / W;  // valid slots of this wavefront in the pass
    // the accumulators start from Q's tiles (clamped indices: every load issues unconditionally)
    d4v acc[kBatch];
    const QPX_GLOBAL_AS double* Qg = as_global(Q);
#pragma unroll
    for (int b = 0; b < kBatch; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 16 * tI[b] + 4 * r + rq, j = 16 * tJ[b] + cq;
        acc[b][r] = Qg[(i < n ? i : n - 1) * n + (j < n ? j : n - 1)];
      }
    __syncthreads();
    for (int e = tid; e < kGtdgRows * (ldl - n); e += NT) {
      const int r = e / (ldl - n);
      chunk[r * ldl + n + (e - r * (ldl - n))] = 0.0;
    }
    // (a full batch, the common case, runs its MFMAs unguarded)
    auto mfmas = [&](auto full) {
#pragma unroll
      for (int ks = 0; ks < kGtdgRows / 4; ++ks) {
        const double* gr = chunk + (4 * ks + rq) * ldl + cq;
        const double dk = cD[4 * ks + rq];
        double a[kBatch], bb[kBatch];
#pragma unroll
        for (int b = 0; b < kBatch; ++b) {
          a[b] = gr[16 * tI[b]] * dk;
          bb[b] = gr[16 * tJ[b]];
        }
#pragma unroll
        for (int b = 0; b < kBatch; ++b)
          if (decltype(full)::value || b < nb) acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[b], bb[b], acc[b], 0, 0, 0);
      }
    };
    auto run_mfmas = [&]() {
      if (nb >= kBatch) mfmas(std::true_type{});
      else mfmas(std::false_type{});
    };
    if constexpr (kPf) {
      // chunk k0 from the registers into LDS, then chunk k0 + 1's loads in flight under its MFMAs
      fetch(0);
      for (int k0 = 0; k0 < mi; k0 += kGtdgRows) {
        __syncthreads();
        const int64_t base = (int64_t)k0 * n;
#pragma unroll
        for (int u = 0; u < EPT; ++u)
          if (dst[u] >= 0) chunk[dst[u]] = base + tid + NT * u <= last ? (double)reg[u] : 0.0;
        if (tid < kGtdgRows) cD[tid] = k0 + tid < mi ? dreg : 0.0;
        __syncthreads();
        if (k0 + kGtdgRows < mi) fetch(k0 + kGtdgRows);
        run_mfmas();
      }
    } else {
      for (int k0 = 0; k0 < mi; k0 += kGtdgRows) {
        __syncthreads();
        for (int r = 0; r < kGtdgRows; ++r) {
          const int k = k0 + r;
          for (int c = tid; c < n; c += NT) chunk[r * ldl + c] = k < mi ? (double)G[(int64_t)k * n + c] : 0.0;
        }
        for (int r = tid; r < kGtdgRows; r += NT) cD[r] = k0 + r < mi ? D[k0 + r] : 0.0;
        __syncthreads();
        run_mfmas();
      }
    }
#pragma unroll
    for (int b = 0; b < kBatch; ++b) {
      if (!ok[b]) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 16 * tI[b] + 4 * r + rq, j = 16 * tJ[b] + cq;
        if (i < n && j <= i) as_global_w(H)[i * n + j] = acc[b][r];
      }
    }
  }
  __syncthreads();
}

// G^T D G for the 256-thread blocks with each wavefront's tiles grouped into 4x4 super-blocks of
// tiles, so that a k-step reads 8 operands from LDS instead of two per tile.  Work units: a full
// super-block (P, Q), P > Q (16 tiles: rows 4P..4P+3 times columns 4Q..4Q+3, A operands the four
// row blocks scaled by D, B operands the four column blocks), or a pair of diagonal super-blocks
// (P1, P1), (P2, P2) (their 10 lower tiles each; a diagonal super-block's four G column blocks
// serve as both operands).  Wavefront w takes units w, w + 4, ... (one pass over G per four
// units: two at n = 256).  Every tile accumulates exactly as in gtdg_mfma (from Q's tile, chunk by
// chunk, k-step by k-step), so H is the same bit for bit.
constexpr int kSbSlots = 20;
__device__ __forceinline__ constexpr int sb_tri_i(int s) { return s < 1 ? 0 : s < 3 ? 1 : s < 6 ? 2 : 3; }
__device__ __forceinline__ constexpr int sb_tri_j(int s) { return s - sb_tri_i(s) * (sb_tri_i(s) + 1) / 2; }

template <typename GT>
__device__ __noinline__ void gtdg_sb(const double* Q, const double* D, double* H, int n, int mi, double* chunk,
                                     const GT* G) {
  constexpr int NT = 256, W = NT / 64, S = kSbSlots;
  constexpr int EPT = (kGtdgRows * (16 * ((kMfmaMaxN / 16) | 1)) + NT - 1) / NT;
  const int tid = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63, rq = l >> 4, cq = l & 15;
  const int nt = (n + 15) / 16, ldl = gtdg_stride(n), ns = (nt + 3) / 4;
  const int nF = ns * (ns - 1) / 2, U = nF + (ns + 1) / 2;
  const int csz = kGtdgRows * n;
  // the chunk keeps G's own element type (widened exactly when read as an operand), then D
  GT* cG = reinterpret_cast<GT*>(chunk);
  double* cD = chunk + (kGtdgRows * ldl * (int)sizeof(GT) + 7) / 8;
  int dst[EPT];
#pragma unroll
  for (int u = 0; u < EPT; ++u) {
    const int e = tid + NT * u;
    dst[u] = e < csz ? (e / n) * ldl + (e - (e / n) * n) : -1;
  }
  GT reg[EPT];
  double dreg = 0.0;
  const QPX_GLOBAL_AS GT* Gg = uniform_global(G);
  const QPX_GLOBAL_AS double* Dg = uniform_global(D);
  const int64_t last = (int64_t)mi * n - 1;
  auto fetch = [&](int k0) {
    const int64_t base = (int64_t)k0 * n;
#pragma unroll
    for (int u = 0; u < EPT; ++u) {
      const int64_t g = base + tid + NT * u;
      reg[u] = Gg[(unsigned)(g < last ? g : last)];  // mi * n < 2^31
    }
    const int kd = k0 + (tid & (kGtdgRows - 1));
    dreg = Dg[(unsigned)(kd < mi ? kd : mi - 1)];
  };
  const QPX_GLOBAL_AS double* Qg = uniform_global(Q);
  QPX_GLOBAL_AS double* Hw = uniform_global_w(H);
  for (int u0 = 0; u0 < U; u0 += W) {
    const int u = u0 + w;
    const bool full = u < nF;  // a full super-block, else a pair of diagonal ones (or nothing)
    int P = 0, Qb = 0;
    if (full) {
      P = (int)((1.0 + sqrt(1.0 + 8.0 * u)) * 0.5);
      while (P * (P - 1) / 2 > u) --P;
      while ((P + 1) * P / 2 <= u) ++P;
      Qb = u - P * (P - 1) / 2;
    } else {
      P = 2 * (u - nF);  // diagonal super-blocks P and P + 1 (Qb unused)
    }
    const bool has = u < U;
    int tI[S], tJ[S];
    bool ok[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      int I, J;
      if (full) {
        I = 4 * P + (s & 15) / 4;
        J = 4 * Qb + (s & 3);
      } else {
        const int sb = s < 10 ? P : P + 1, t = s < 10 ? s : s - 10;
        I = 4 * sb + sb_tri_i(t);
        J = 4 * sb + sb_tri_j(t);
      }
      ok[s] = has && (full ? s < 16 : true) && I < nt && J < nt;
      tI[s] = ok[s] ? I : 0;
      tJ[s] = ok[s] ? J : 0;
    }
    d4v acc[S];
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 16 * tI[s] + 4 * r + rq, j = 16 * tJ[s] + cq;
        acc[s][r] = Qg[(unsigned)((i < n ? i : n - 1) * n + (j < n ? j : n - 1))];
      }
    __syncthreads();
    for (int e = tid; e < kGtdgRows * (ldl - n); e += NT) {
      const int r = e / (ldl - n);
      cG[r * ldl + n + (e - r * (ldl - n))] = (GT)0;
    }
    fetch(0);
    for (int k0 = 0; k0 < mi; k0 += kGtdgRows) {
      __syncthreads();
      {
        const int64_t base = (int64_t)k0 * n;
#pragma unroll
        for (int u = 0; u < EPT; ++u)
          if (dst[u] >= 0) cG[dst[u]] = base + tid + NT * u <= last ? reg[u] : (GT)0;
        if (tid < kGtdgRows) cD[tid] = k0 + tid < mi ? dreg : 0.0;
      }
      __syncthreads();
      if (k0 + kGtdgRows < mi) fetch(k0 + kGtdgRows);
      if (!has) continue;
      if (full) {
#pragma unroll
        for (int ks = 0; ks < kGtdgRows / 4; ++ks) {
          const GT* gr = cG + (4 * ks + rq) * ldl + cq;
          const GT* ga = gr + 64 * P;
          const GT* gb = gr + 64 * Qb;
          const double dk = cD[4 * ks + rq];
          double a[4], b[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            a[i] = (double)ga[16 * i] * dk;
            b[i] = (double)gb[16 * i];
          }
#pragma unroll
          for (int s = 0; s < 16; ++s)
            acc[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s / 4], b[s & 3], acc[s], 0, 0, 0);
        }
      } else {
#pragma unroll
        for (int ks = 0; ks < kGtdgRows / 4; ++ks) {
          const GT* gr = cG + (4 * ks + rq) * ldl + cq;
          const GT* g1 = gr + 64 * P;
          const double dk = cD[4 * ks + rq];
          double b1[8], a1[8];
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            b1[i] = (double)g1[16 * i];  // column blocks 4P .. 4P + 7: both diagonal super-blocks
            a1[i] = b1[i] * dk;
          }
#pragma unroll
          for (int s = 0; s < S; ++s) {
            const int o = s < 10 ? 0 : 4, t = s < 10 ? s : s - 10;
            acc[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[o + sb_tri_i(t)], b1[o + sb_tri_j(t)], acc[s], 0, 0, 0);
          }
        }
      }
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (!ok[s]) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 16 * tI[s] + 4 * r + rq, j = 16 * tJ[s] + cq;
        if (i < n && j <= i) Hw[(unsigned)(i * n + j)] = acc[s][r];
      }
    }
  }
  __syncthreads();
}

// block_cholesky for m <= kSmallChol (the Schur complement F = A H^{-1} A^T) on one wavefront:
// lane r holds row r, and each column reaches every lane through a 32-double LDS broadcast
// instead of global round trips and two workgroup barriers per column.  Same arithmetic as
// block_cholesky (IEEE square root and division, the same products in the same order), so L is
// the same bit for bit.  S: (m x 33 + 32) doubles of free LDS.
constexpr int kSmallChol = 32;
template <int NT>
__device__ void block_cholesky_small(double* M_, int m, double* diag0_, double* rcp_, double* S_) {
  QPX_ASSUME_LDS(S_);
  QPX_LDS_AS double* S = (QPX_LDS_AS double*)S_;
  QPX_LDS_AS double* colb = S + kSmallChol * 33;
  QPX_GLOBAL_AS double* M = as_global_w(M_);
  QPX_GLOBAL_AS double* diag0 = as_global_w(diag0_);
  QPX_GLOBAL_AS double* rcp = as_global_w(rcp_);
  const int tid = threadIdx.x, l = tid & 63;
  __syncthreads();
  for (int e = tid; e < m * m; e += NT) {
    const int r = e / m, c = e - r * m;
    S[r * 33 + c] = M[e];
  }
  __syncthreads();
  if (tid < 64) {
    const int r = l < m ? l : m - 1;
    double p[kSmallChol];
#pragma unroll
    for (int c = 0; c < kSmallChol; ++c) p[c] = c < m ? S[r * 33 + c] : 0.0;
    if (l < m) diag0[l] = S[l * 33 + l];
#pragma unroll
    for (int j = 0; j < kSmallChol; ++j) {
      if (j < m) {
        if (l < m) colb[l] = p[j];
        __builtin_amdgcn_wave_barrier();
        double a[kSmallChol];
#pragma unroll
        for (int c = j; c < kSmallChol; ++c) a[c] = c < m ? colb[c] : 0.0;
        __builtin_amdgcn_wave_barrier();
        const double d = a[j];
        const bool ok = d > kPivotRel * fabs(S[j * 33 + j]);
        const double ljj = ok ? sqrt(d) : kHugePivot;
        const double inv = 1.0 / ljj;
        if (r > j) p[j] *= inv;
        if (r == j) p[j] = ljj;
        if (l == 0) rcp[j] = inv;
#pragma unroll
        for (int c = j + 1; c < kSmallChol; ++c) {
          const double lcj = a[c] * inv;
          if (c < m && r >= c) p[c] -= p[j] * lcj;
        }
      }
    }
    if (l < m)
#pragma unroll
      for (int c = 0; c < kSmallChol; ++c)
        if (c < m) M[l * m + c] = p[c];
  }
  __syncthreads();
}

// Pieces of the blocked Cholesky (64 <= m <= kMfmaMaxN) on block column k, held as a panel in LDS:
// rows c0..m-1 of columns c0..c0+cw-1 at row stride kPanelLd, zero above the diagonal and beyond cw
// columns.

// The 16x16 diagonal block on wavefront 0, lane r holding row r (right-looking, as the scalar
// block_cholesky: scale column j by 1/L_jj, then subtract its outer product); 1/L_jj to rcp and
// rcpl (LDS).  Per column j, lanes 0..15 publish their unscaled entry of column j in colb (16
// doubles of LDS) and read the whole column back, so every lane forms L_jj, 1/L_jj and the
// multipliers a_cj / L_jj itself: one LDS write and read per column, no readlane chain.  The pivot
// test's reference diagonal (diag0, in LDS) is read per column beside the column.  1/L_jj
// comes from v_rsq_f64 and two Newton steps (L_jj = d / L_jj): seven dependent operations on the
// column's chain instead of a correctly rounded square root and division (about 22), within an
// ulp or two of the rounded values.  A pivot must also be a normal number (d >= 2^-1022) for the
// seed; a denormal one is replaced like one below the relative floor.  FULL: cw == 16 at compile
// time, so the unrolled columns carry no width tests.  The block's inverse goes to linv (16 rows
// at stride kPanelLd, zero past cw): lane r applies each column's scaling and eliminations, with
// the multipliers it already forms, to column r of the identity, off the pivot chain.
template <bool FULL>
__device__ __forceinline__ void panel_diag_cols(double* cur_, int c0, int cw, const double* diag0_, double* rcp_,
                                                double* rcpl_, double* colb_, double* linv_) {
  QPX_LDS_AS double* cur = (QPX_LDS_AS double*)cur_;
  QPX_LDS_AS double* rcpl = (QPX_LDS_AS double*)rcpl_;
  QPX_LDS_AS double* colb = (QPX_LDS_AS double*)colb_;
  const QPX_LDS_AS double* diag0 = (const QPX_LDS_AS double*)diag0_;
  QPX_GLOBAL_AS double* rcp = as_global_w(rcp_);
  QPX_LDS_AS double* linv = (QPX_LDS_AS double*)linv_;
  const int l = threadIdx.x & 63, r = l & 15;
  if (FULL) cw = 16;
  // lane r also holds column r of L^{-1}, built by the same eliminations (identity at the start)
  double p[16], e[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) e[c] = c == r ? 1.0 : 0.0;
#pragma unroll
  for (int c = 0; c < 16; ++c) p[c] = cur[r * kPanelLd + c];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    if (FULL || j < cw) {
      if (l < 16) colb[r] = p[j];
      __builtin_amdgcn_wave_barrier();
      double a[16];
#pragma unroll
      for (int c = j; c < 16; ++c) a[c] = colb[c];
      __builtin_amdgcn_wave_barrier();
      const double d = a[j];
      // the reference diagonal from LDS (a broadcast read beside the column's)
      const bool ok = d > kPivotRel * fabs(diag0[c0 + j]) && d >= 0x1p-1022;
      // the seed and the Newton steps run on d unconditionally (whatever they give for a
      // rejected pivot is discarded by the select)
      const double h = 0.5 * d;
      double y = __builtin_amdgcn_rsq(d);
      y = fma(y, fma(-(h * y), y, 0.5), y);
      y = fma(y, fma(-(h * y), y, 0.5), y);
      const double inv = ok ? y : 1.0 / kHugePivot, ljj = ok ? d * y : kHugePivot;
      if (r > j) p[j] *= inv;
      if (r == j) p[j] = ljj;
      if (l == 0) {
        rcp[c0 + j] = inv;
        rcpl[c0 + j] = inv;
      }
      e[j] *= inv;
#pragma unroll
      for (int c = j + 1; c < 16; ++c) {
        const double lcj = a[c] * inv;
        if (r >= c && (FULL || c < cw)) p[c] -= p[j] * lcj;
        e[c] -= lcj * e[j];
      }
    }
  }
  if (l < 16) {
#pragma unroll
    for (int c = 0; c < 16; ++c) cur[r * kPanelLd + c] = p[c];
    // L^{-1}[j][r] (zero past the block's width)
#pragma unroll
    for (int j = 0; j < 16; ++j) linv[j * kPanelLd + r] = (FULL || (j < cw && r < cw)) ? e[j] : 0.0;
  }
}
__device__ __forceinline__ void panel_diag_lds(double* cur_, int c0, int cw, const double* diag0_, double* rcp_,
                                               double* rcpl_, double* colb_, double* linv_) {
  if (cw == 16) panel_diag_cols<true>(cur_, c0, cw, diag0_, rcp_, rcpl_, colb_, linv_);
  else panel_diag_cols<false>(cur_, c0, cw, diag0_, rcp_, rcpl_, colb_, linv_);
}

// The panel's rows below the diagonal block, L[r, :] = P[r, :] L_kk^{-T}, and with Wl W's rows of
// the block, W[c0.., :] = L_kk^{-1} W[c0.., :], as 16x16 tiles of four v_mfma_f64_16x16x4f64
// each, in place: items (W's column tiles first, then the row tiles) go to wavefronts wi, wi + ws,
// ...  Each tile reads its operands into registers before it writes, and no two tiles share
// entries.
__device__ __forceinline__ void panel_rows_mfma(double* cur_, int R, int c0, int cw, const double* linv_,
                                                double* Wl_, int me, int wi, int ws) {
  QPX_LDS_AS double* cur = (QPX_LDS_AS double*)cur_;
  const QPX_LDS_AS double* linv = (const QPX_LDS_AS double*)linv_;
  QPX_LDS_AS double* Wl = (QPX_LDS_AS double*)Wl_;
  const int l = threadIdx.x & 63, rq = l >> 4, cq = l & 15;
  const int nw = Wl_ != nullptr ? (me + 15) / 16 : 0, nr = (R + 15) / 16 - 1;
  for (int it = wi; it < nw + nr; it += ws) {
    double av[4], bv[4];
    d4v acc = {0.0, 0.0, 0.0, 0.0};
    if (it < nw) {  // W's columns 16 it .. 16 it + 15: L_kk^{-1} (A) times W's block rows (B)
      const int col = 16 * it + cq, colc = col < me ? col : me - 1;
#pragma unroll
      for (int sl = 0; sl < 4; ++sl) {
        const int j = 4 * sl + rq;
        av[sl] = linv[cq * kPanelLd + j];
        const double wv = Wl[(c0 + (j < cw ? j : cw - 1)) * me + colc];
        bv[sl] = j < cw ? wv : 0.0;
      }
#pragma unroll
      for (int sl = 0; sl < 4; ++sl) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[sl], bv[sl], acc, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 4 * r + rq;
        if (i < cw && col < me) Wl[(c0 + i) * me + col] = acc[r];
      }
    } else {  // panel rows 16 t .. 16 t + 15 (A) times L_kk^{-T} (B)
      const int t = it - nw + 1;
#pragma unroll
      for (int sl = 0; sl < 4; ++sl) {
        av[sl] = cur[(16 * t + cq) * kPanelLd + 4 * sl + rq];
        bv[sl] = linv[cq * kPanelLd + 4 * sl + rq];
      }
#pragma unroll
      for (int sl = 0; sl < 4; ++sl) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[sl], bv[sl], acc, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = 16 * t + 4 * r + rq;
        if (row < R) cur[row * kPanelLd + cq] = cq < cw ? acc[r] : 0.0;
      }
    }
  }
}

template <int NT>
__device__ __forceinline__ void panel_store(double* M_, int m, const double* cur_, int R, int c0, int cw, int t0 = 0) {
  QPX_GLOBAL_AS double* M = uniform_global_w(M_);
  const QPX_LDS_AS double* cur = (const QPX_LDS_AS double*)cur_;
  for (int e = (int)threadIdx.x - t0; e < R * 16; e += NT - t0) {
    const int r = e >> 4, c = e & 15;
    if (c < cw && (r >= 16 || c <= r)) M[(unsigned)((c0 + r) * m + c0 + c)] = cur[r * kPanelLd + c];
  }
}


// W's rows below block k receive -L[rows, block] W[block, :] as 16x16 tiles (block_w_lds's order),
// over wavefronts w0..W-1.
template <int NT>
__device__ __forceinline__ void panel_w_trailing(double* Wl_, const double* cur_, int k, int m, int me, int w0 = 0) {
  QPX_LDS_AS double* Wl = (QPX_LDS_AS double*)Wl_;
  const QPX_LDS_AS double* cur = (const QPX_LDS_AS double*)cur_;
  constexpr int W = NT / 64;
  const int tid = threadIdx.x, w = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63, rq = l >> 4, cq = l & 15;
  const int nt = (m + 15) / 16, ct = (me + 15) / 16, c0 = 16 * k, cw = m - c0 < 16 ? m - c0 : 16;
  const int Tw = (nt - k - 1) * ct;
  for (int idx = w - w0; idx < Tw; idx += W - w0) {
    const int Rb = k + 1 + idx / ct, C = idx - (idx / ct) * ct;
    const int ia = 16 * Rb + cq, jb = 16 * C + cq;
    const int pa = (ia < m ? ia : m - 1) - c0, jbc = jb < me ? jb : me - 1;
    double av[4], bv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int kk = 4 * q + rq;
      const int kc = kk < cw ? kk : cw - 1;
      const double lv = cur[pa * kPanelLd + kc], wv = Wl[(c0 + kc) * me + jbc];
      av[q] = (ia < m && kk < cw) ? -lv : 0.0;
      bv[q] = (kk < cw && jb < me) ? wv : 0.0;
    }
    d4v wacc;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = 16 * Rb + 4 * r + rq;
      const double wv = Wl[(i < m ? i : m - 1) * me + jbc];
      wacc[r] = (i < m && jb < me) ? wv : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) wacc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q], bv[q], wacc, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = 16 * Rb + 4 * r + rq;
      if (i < m && jb < me) Wl[i * me + jb] = wacc[r];
    }
  }
}

// Blocked Cholesky for 64 <= m <= kMfmaMaxN, same contract as block_cholesky except that the
// reference diagonal of the pivot test stays in LDS (no diag0).  Block column k (16 columns, rows
// 16k..m-1) is a panel in LDS: its 16x16 diagonal block is factored in one wavefront's registers
// and inverted (L_kk^{-1}), the rows below become P L_kk^{-T} as MFMA tiles, and the panel is
// written back.  Trailing tiles (I, J) receive -L_Ik L_Jk^T as four v_mfma_f64_16x16x4f64 per
// block column, two block columns per pass (the step order is spelled out in the body),
// kTrailGroup tiles at a time per wavefront with the next group's loads in flight.  With Wl
// (n x me in LDS beside the panels, holding A^T on entry), W = L^{-1} A^T is formed alongside:
// W's rows of each block become L_kk^{-1} times themselves (MFMA tiles beside the panel's rows),
// and the rows below take the block's update.  P: two panels of panel_words(m); aux:
// 2 m + 16 + 16 kPanelLd doubles of LDS for 1/L_jj, the diagonal block's column broadcasts, the
// reference diagonal of the pivot test and L_kk^{-1}.
constexpr int kTrailGroup = 4;  // trailing tiles per group: the next group's loads fly under its MFMAs
template <int NT>
__device__ __noinline__ void block_cholesky_tiled(double* M, int m, double* rcp, double* P,
                                                  double* Wl, int me, double* aux) {
  constexpr int W = NT / 64;
  const int tid = threadIdx.x, w = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63, rq = l >> 4, cq = l & 15;
  const int nt = (m + 15) / 16;
  const bool has_w = Wl != nullptr;
  if (!has_w) Wl = aux;  // never dereferenced then; an LDS address, so the assumption below holds
  QPX_ASSUME_LDS(P);
  QPX_ASSUME_LDS(aux);
  QPX_ASSUME_LDS(Wl);
  const QPX_GLOBAL_AS double* Mg = uniform_global(M);
  QPX_GLOBAL_AS double* Mw = uniform_global_w(M);
  double* cur = P;
  double* nxt = P + panel_words(m);
  // the reference diagonal of the pivot test, in LDS past 1/L_jj and the column buffer
  double* dl = aux + m + 16;
  // L_kk^{-1} of the latest factored diagonal block (16 rows at stride kPanelLd) past it
  double* linv = dl + m;
  for (int i = tid; i < m; i += NT) dl[i] = Mg[(unsigned)(i * m + i)];
  // rows r0.. of block column kb from M into a panel (zero above the diagonal and beyond the block
  // column's width), eight loads in flight per thread through clamped indices
  auto load_panel = [&](double* pnl_, int kb, int r0) {
    QPX_LDS_AS double* pnl = (QPX_LDS_AS double*)pnl_;
    const int c0 = 16 * kb, R = m - c0, cw = R < 16 ? R : 16;
    for (int e0 = 16 * r0; e0 < R * 16; e0 += 8 * NT) {
      double val[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + tid + u * NT, r = e >> 4 < R ? e >> 4 : R - 1, c = (e & 15) < cw ? e & 15 : cw - 1;
        val[u] = Mg[(unsigned)((c0 + r) * m + c0 + c)];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + tid + u * NT, r = e >> 4, c = e & 15;
        if (r < R) pnl[r * kPanelLd + c] = (c < cw && (r >= 16 || c <= r)) ? val[u] : 0.0;
      }
    }
  };
  load_panel(cur, 0, 0);
  __syncthreads();
  if (w == 0) {
    panel_diag_lds(cur, 0, m < 16 ? m : 16, dl, rcp, aux, aux + m, linv);
  }
  __syncthreads();
  auto load = [&](int I, int J, d4v& acc) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = 16 * I + 4 * r + rq, jj = 16 * J + cq;
      acc[r] = Mg[(unsigned)((i < m ? i : m - 1) * m + (jj < m ? jj : m - 1))];
    }
  };
  // acc -= L_Ik L_Jk^T from the panel of block column k (rows past m count as zero)
  auto update = [&](const double* pnl, int k, int R, int I, int J, d4v& acc) {
    const int ia = 16 * (I - k) + cq, ib = 16 * (J - k) + cq;
    const double* pa = pnl + ia * kPanelLd + rq;
    const double* pb = pnl + ib * kPanelLd + rq;
#pragma unroll
    for (int sl = 0; sl < 4; ++sl) {
      const double pav = pa[4 * sl], pbv = pb[4 * sl];
      const double av = ia < R ? -pav : 0.0, bv = ib < R ? pbv : 0.0;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
    }
  };
  // the trailing tiles (first + a, first + b), b <= a < tn2, idx >= i0 in row-major order, receive
  // the updates of block column k (panel pk) and, with pk1, of block column k + 1 (panel pk1), in
  // groups of kTrailGroup consecutive tiles that the wavefronts take from an LDS counter (so a
  // wavefront that comes late takes fewer), the next group loaded before the current group's MFMAs
  __shared__ int trail_next;
  auto trail = [&](const double* pk, const double* pk1, int k, int first, int tn2, int i0) {
    constexpr int GS = kTrailGroup;
    const int T = tn2 * (tn2 + 1) / 2, R = m - 16 * k, ng = (T - i0 + GS - 1) / GS;
    auto grab = [&]() {
      int g = 0;
      if (l == 0) g = atomicAdd(&trail_next, 1);
      return __builtin_amdgcn_readfirstlane(g);
    };
    auto decode = [&](int g, int t, int& I, int& J) {
      const int idx = i0 + GS * g + t < T ? i0 + GS * g + t : T - 1;
      int a = (int)((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
      while ((a + 1) * (a + 2) / 2 <= idx) ++a;
      while (a * (a + 1) / 2 > idx) --a;
      I = first + a;
      J = first + (idx - a * (a + 1) / 2);
    };
    int g = grab();
    if (g >= ng) return;
    int tI[GS] = {}, tJ[GS] = {}, uI[GS] = {}, uJ[GS] = {};
    d4v acc[GS] = {}, nacc[GS] = {};
#pragma unroll
    for (int t = 0; t < GS; ++t) { decode(g, t, tI[t], tJ[t]); load(tI[t], tJ[t], acc[t]); }
    while (true) {
      const int gn = grab();
      if (gn < ng) {
#pragma unroll
        for (int t = 0; t < GS; ++t) { decode(gn, t, uI[t], uJ[t]); load(uI[t], uJ[t], nacc[t]); }
      }
#pragma unroll
      for (int t = 0; t < GS; ++t) {
        update(pk, k, R, tI[t], tJ[t], acc[t]);
        if (pk1 != nullptr) update(pk1, k + 1, R - 16, tI[t], tJ[t], acc[t]);
      }
      // every group stores the same number of entries, so that the next group's loads can be
      // awaited without waiting for these stores: entries outside the matrix or past the last
      // tile go to M[0][m - 1], in the strict upper triangle, which nothing reads
#pragma unroll
      for (int t = 0; t < GS; ++t) {
        const bool valid = i0 + GS * g + t < T;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = 16 * tI[t] + 4 * r + rq, jj = 16 * tJ[t] + cq;
          Mw[(unsigned)(valid && i < m && jj < m ? i * m + jj : m - 1)] = acc[t][r];
        }
      }
      if (gn >= ng) break;
#pragma unroll
      for (int t = 0; t < GS; ++t) { acc[t] = nacc[t]; tI[t] = uI[t]; tJ[t] = uJ[t]; }
      g = gn;
    }
  };
  // the first wavefront factors diagonal blocks, the others (all of them with one wavefront) the
  // rest of each phase
  const int w0 = W > 1 ? 1 : 0;
  const bool diag_wave = w == 0, work_wave = W == 1 || w > 0;
  // Block columns in pairs (k, k + 1) with A = panel k (diagonal block factored) on entry:
  //   1. A's rows below the diagonal block (W's rows of block k beside them); store A
  //   2. block column k + 1 from M with block column k's update into B; its diagonal block factored
  //      while the others apply block column k to W's rows below it
  //   3. B's rows below the diagonal block (W's rows of block k + 1); then, while the first
  //      wavefront forms the diagonal block of block column k + 2 from both panels into A's first
  //      16 rows (free once A is stored; the trailing update reads A's rows from 32 and B's from
  //      16) and factors it, the others store B and start giving every other trailing tile both
  //      panels' updates in one pass through M (the first wavefront joins when its block is done);
  //      then all apply block column k + 1 to W's rows below it
  //   4. the rows of block column k + 2 below its diagonal block from M into A
  // Each tile sees the updates of block columns 0, 1, ... in that order, as with one block column
  // per step, and the trailing matrix goes through M once per pair.
  double* A = cur;
  double* B = nxt;
  auto rows_of = [&](double* pnl, int kb) {
    const int c0 = 16 * kb, R = m - c0, cw = R < 16 ? R : 16;
    panel_rows_mfma(pnl, R, c0, cw, linv, has_w ? Wl : nullptr, me, w, W);
  };
  int k = 0;
  for (; k + 1 < nt; k += 2) {
    const int c0 = 16 * k, R = m - c0, cw = 16;
    const int cw1 = R - 16 < 16 ? R - 16 : 16, tn = nt - k - 1;
    rows_of(A, k);
    __syncthreads();
    panel_store<NT>(M, m, A, R, c0, cw);
    for (int a = w; a < tn; a += W) {
      const int I = k + 1 + a;
      d4v acc;
      load(I, k + 1, acc);
      update(A, k, R, I, k + 1, acc);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int pr_ = 16 * a + 4 * r + rq;
        if (16 * I + 4 * r + rq < m) B[pr_ * kPanelLd + cq] = (cq < cw1 && (pr_ >= 16 || cq <= pr_)) ? acc[r] : 0.0;
      }
    }
    __syncthreads();
    if (diag_wave) {
      panel_diag_lds(B, c0 + 16, cw1, dl, rcp, aux, aux + m, linv);
    }
    if (work_wave && has_w) panel_w_trailing<NT>(Wl, A, k, m, me, w0);
    __syncthreads();
    rows_of(B, k + 1);
    if (tid == 0) trail_next = 0;
    __syncthreads();
    const bool more = k + 2 < nt;
    const int cw2 = more ? (R - 32 < 16 ? R - 32 : 16) : 0;
    if (diag_wave && more) {
      d4v acc;
      load(k + 2, k + 2, acc);
      update(A, k, R, k + 2, k + 2, acc);
      update(B, k + 1, R - 16, k + 2, k + 2, acc);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int pr_ = 4 * r + rq;
        if (c0 + 32 + pr_ < m) A[pr_ * kPanelLd + cq] = (cq < cw2 && cq <= pr_) ? acc[r] : 0.0;
      }
      __builtin_amdgcn_wave_barrier();
      panel_diag_lds(A, c0 + 32, cw2, dl, rcp, aux, aux + m, linv);
    }
    if (work_wave) panel_store<NT>(M, m, B, R - 16, c0 + 16, cw1, 64 * w0);
    if (more) trail(A, B, k, k + 2, tn - 1, 1);
    if (has_w) panel_w_trailing<NT>(Wl, B, k + 1, m, me, 0);
    __syncthreads();
    if (more) {
      load_panel(A, k + 2, 16);
      __syncthreads();
    }
  }
  if (k < nt) {  // an odd number of block columns: the last one alone
    const int c0 = 16 * k, R = m - c0, cw = R < 16 ? R : 16;
    rows_of(A, k);
    __syncthreads();
    panel_store<NT>(M, m, A, R, c0, cw);
    __syncthreads();
  }
}

// Blocked forward substitution L y = b (L from block_cholesky[_tiled]; b overwritten): wavefront 0
// eliminates inside each 16-row block with lane broadcasts, then all threads apply the block to
// the rows below it.
template <int NT>
__device__ __forceinline__ void fsub_blocks(const double* L, const double* rcp, int m, double* b) {
  const int tid = threadIdx.x, l = tid & 63;
  const int nt = (m + 15) / 16;
  const QPX_GLOBAL_AS double* Lg = as_global(L);
  const QPX_GLOBAL_AS double* rg = as_global(rcp);
  for (int k = 0; k < nt; ++k) {
    const int c0 = 16 * k, cw = m - c0 < 16 ? m - c0 : 16;
    if (tid < 64) {
      // row l of the diagonal block and the block's 1/L_jj, loaded up front (clamped indices)
      const int lr = c0 + (l < cw ? l : cw - 1);
      double lrow[16], rc[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int jc = c0 + (j < cw ? j : cw - 1);
        lrow[j] = Lg[lr * m + jc];
        rc[j] = rg[jc];
      }
      double bi = l < cw ? b[c0 + l] : 0.0;
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        if (j < cw) {
          const double yj = readlane_d(bi, j) * rc[j];
          const double lv = (l > j && l < cw) ? lrow[j] : 0.0;
          bi = (l == j) ? yj : bi - lv * yj;
        }
      }
      if (l < cw) b[c0 + l] = bi;
    }
    __syncthreads();
    for (int r = c0 + 16 + tid; r < m; r += NT) {
      double lv[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) lv[j] = Lg[r * m + c0 + (j < cw ? j : cw - 1)];
      double a0 = 0.0, a1 = 0.0;
#pragma unroll
      for (int j = 0; j < 16; j += 2) {
        if (j < cw) a0 += lv[j] * b[c0 + j];
        if (j + 1 < cw) a1 += lv[j + 1] * b[c0 + j + 1];
      }
      b[r] -= a0 + a1;
    }
    __syncthreads();
  }
}

// Blocked back substitution L^T x = b (b overwritten), blocks from the last.
template <int NT>
__device__ __forceinline__ void bsub_t_blocks(const double* L, const double* rcp, int m, double* b) {
  const int tid = threadIdx.x, l = tid & 63;
  const int nt = (m + 15) / 16;
  const QPX_GLOBAL_AS double* Lg = as_global(L);
  const QPX_GLOBAL_AS double* rg = as_global(rcp);
  for (int k = nt - 1; k >= 0; --k) {
    const int c0 = 16 * k, cw = m - c0 < 16 ? m - c0 : 16;
    if (tid < 64) {
      // column l of the diagonal block and the block's 1/L_jj, loaded up front (clamped indices)
      const int lc = c0 + (l < cw ? l : cw - 1);
      double lcol[16], rc[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int jc = c0 + (j < cw ? j : cw - 1);
        lcol[j] = Lg[jc * m + lc];
        rc[j] = rg[jc];
      }
      double bi = l < cw ? b[c0 + l] : 0.0;
#pragma unroll
      for (int j = 15; j >= 0; --j) {
        if (j < cw) {
          const double xj = readlane_d(bi, j) * rc[j];
          const double lv = (l < j) ? lcol[j] : 0.0;
          bi = (l == j) ? xj : bi - lv * xj;
        }
      }
      if (l < cw) b[c0 + l] = bi;
    }
    __syncthreads();
    for (int r = tid; r < c0; r += NT) {
      double lv[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) lv[j] = Lg[(c0 + (j < cw ? j : cw - 1)) * m + r];
      double a0 = 0.0, a1 = 0.0;
#pragma unroll
      for (int j = 0; j < 16; j += 2) {
        if (j < cw) a0 += lv[j] * b[c0 + j];
        if (j + 1 < cw) a1 += lv[j + 1] * b[c0 + j + 1];
      }
      b[r] -= a0 + a1;
    }
    __syncthreads();
  }
}

// Forward substitution L y = b with b in LDS and one barrier per 16-row block: at step k,
// wavefront 0 gives block k's rows block k - 1's update (y_{k-1} is final) and eliminates inside
// the block, while the other wavefronts apply block k - 1 to the rows from block k + 1 on (at most
// two rows per thread: m - 16 <= 2 (NT - 64)).  Each thread's entries of L for step k + 2 are
// loaded while step k runs: on wavefront 0, lane l's row c0 + l from column c0 - 16 and 1/L_ll;
// on the others, their two rows' 16 entries of block column k - 1.  Every row sees the blocks'
// updates in the same order and with the same arithmetic as fsub_blocks.
template <int NT>
__device__ __noinline__ void fsub_lds(const double* L, const double* rcp, int m, double* b_) {
  constexpr int NW = NT - 64;
  QPX_LDS_AS double* b = (QPX_LDS_AS double*)b_;
  const int tid = threadIdx.x, w = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63, t = tid - 64;
  const int nt = (m + 15) / 16;
  const QPX_GLOBAL_AS double* Lg = uniform_global(L);
  const QPX_GLOBAL_AS double* rg = uniform_global(rcp);
  // a block with a predecessor and all 16 columns inside the matrix reads each row as one run
  // (immediate offsets from one address); the first and a narrow last block clamp every index
  auto fetch = [&](int k, double (&v)[32], double& rl) {
    if (k >= nt) return;
    const int c0 = 16 * k;
    const bool run = k > 0 && c0 + 16 <= m;
    if (w == 0) {
      const int cw = m - c0 < 16 ? m - c0 : 16, r = c0 + (l < cw ? l : cw - 1);
      if (run) {
        const QPX_GLOBAL_AS double* rp = Lg + (unsigned)(r * m + c0 - 16);
#pragma unroll
        for (int j = 0; j < 32; ++j) v[j] = rp[j];
      } else {
#pragma unroll
        for (int j = 0; j < 32; ++j) {
          const int c = c0 - 16 + j, cc = c < 0 ? 0 : (c < m ? c : m - 1);
          v[j] = Lg[(unsigned)(r * m + cc)];
        }
      }
      rl = rg[(unsigned)r];
    } else {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int r = 16 * (k + 1) + t + h * NW, rr = r < m ? r : m - 1;
        const QPX_GLOBAL_AS double* rp = Lg + (unsigned)(rr * m + (k > 0 ? 16 * (k - 1) : 0));
#pragma unroll
        for (int j = 0; j < 16; ++j) v[16 * h + j] = rp[j];
      }
    }
  };
  // FULL: block k is 16 rows wide (every block but the last), so the elimination runs with no
  // width tests (they would make each column a branch that waits for every load in flight)
  auto step = [&](int k, double (&v)[32], double& rl, auto full) {
    constexpr bool FULL = decltype(full)::value;
    const int c0 = 16 * k, cw = FULL ? 16 : (m - c0 < 16 ? m - c0 : 16);
    if (w == 0) {
      double bi = (FULL || l < cw) ? b[c0 + l] : 0.0;
      if (k > 0) {
        double a0 = 0.0, a1 = 0.0;
#pragma unroll
        for (int j = 0; j < 16; j += 2) {
          a0 += v[j] * b[c0 - 16 + j];
          a1 += v[j + 1] * b[c0 - 16 + j + 1];
        }
        bi -= a0 + a1;
      }
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        if (FULL || j < cw) {
          const double yj = readlane_d(bi * rl, j);
          const double lv = (l > j && (FULL || l < cw)) ? v[16 + j] : 0.0;
          bi = (l == j) ? yj : bi - lv * yj;
        }
      }
      if (l < cw) b[c0 + l] = bi;
    } else if (k > 0) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int r = 16 * (k + 1) + t + h * NW;
        if (r < m) {
          double a0 = 0.0, a1 = 0.0;
#pragma unroll
          for (int j = 0; j < 16; j += 2) {
            a0 += v[16 * h + j] * b[c0 - 16 + j];
            a1 += v[16 * h + j + 1] * b[c0 - 16 + j + 1];
          }
          b[r] -= a0 + a1;
        }
      }
    }
    fetch(k + 2, v, rl);
    __syncthreads();
  };
  double va[32], vb[32], ra = 0.0, rb = 0.0;
  fetch(0, va, ra);
  fetch(1, vb, rb);
  const std::integral_constant<bool, true> full;
  const std::integral_constant<bool, false> part;
  int k = 0;
  for (; k + 2 < nt; k += 2) {  // blocks k and k + 1 both full
    step(k, va, ra, full);
    step(k + 1, vb, rb, full);
  }
  if (k < nt) {
    if (k + 1 < nt) {
      step(k, va, ra, full);
      step(k + 1, vb, rb, part);
    } else {
      step(k, va, ra, part);
    }
  }
}

// Back substitution L^T x = b with b in LDS, blocks from the last, one barrier per block: at step
// k, wavefront 0 gives block k's rows block k + 1's update and eliminates inside the block (lane l
// holds column c0 + l of rows c0..c0 + 31), while the other wavefronts apply block k + 1 to the rows
// above block k.  Same order and arithmetic per row as bsub_t_blocks; m - 16 <= 2 (NT - 64).
template <int NT>
__device__ __noinline__ void bsub_t_lds(const double* L, const double* rcp, int m, double* b_) {
  constexpr int NW = NT - 64;
  QPX_LDS_AS double* b = (QPX_LDS_AS double*)b_;
  const int tid = threadIdx.x, w = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63, t = tid - 64;
  const int nt = (m + 15) / 16;
  const QPX_GLOBAL_AS double* Lg = uniform_global(L);
  const QPX_GLOBAL_AS double* rg = uniform_global(rcp);
  // the row index of every load is the same on all lanes (a scalar base), the column the per-lane
  // offset
  auto fetch = [&](int k, double (&v)[32], double& rl) {
    if (k < 0) return;
    const int c0 = 16 * k;
    if (w == 0) {
      const int cw = m - c0 < 16 ? m - c0 : 16, c = c0 + (l < cw ? l : cw - 1);
#pragma unroll
      for (int j = 0; j < 32; ++j) {
        const int r = c0 + j;
        v[j] = (Lg + (unsigned)((r < m ? r : m - 1) * m))[(unsigned)c];
      }
      rl = rg[(unsigned)c];
    } else {
      const int c1 = c0 + 16;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int r = t + h * NW, rr = r < c0 ? r : 0;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          const int i = c1 + j;
          v[16 * h + j] = (Lg + (unsigned)((i < m ? i : m - 1) * m))[(unsigned)rr];
        }
      }
    }
  };
  // FULL: blocks k and k + 1 are both 16 rows wide (k < nt - 2): no width tests
  auto step = [&](int k, double (&v)[32], double& rl, auto full) {
    constexpr bool FULL = decltype(full)::value;
    const int c0 = 16 * k, cw = FULL ? 16 : (m - c0 < 16 ? m - c0 : 16);
    const int c1 = c0 + 16, cw1 = FULL ? 16 : (m - c1 < 16 ? m - c1 : 16);  // block k + 1 (when k + 1 < nt)
    if (w == 0) {
      double bi = (FULL || l < cw) ? b[c0 + l] : 0.0;
      if (k + 1 < nt) {
        double a0 = 0.0, a1 = 0.0;
#pragma unroll
        for (int j = 0; j < 16; j += 2) {
          if (FULL || j < cw1) a0 += v[16 + j] * b[c1 + j];
          if (FULL || j + 1 < cw1) a1 += v[16 + j + 1] * b[c1 + j + 1];
        }
        bi -= a0 + a1;
      }
#pragma unroll
      for (int j = 15; j >= 0; --j) {
        if (FULL || j < cw) {
          const double xj = readlane_d(bi * rl, j);
          const double lv = (l < j) ? v[j] : 0.0;
          bi = (l == j) ? xj : bi - lv * xj;
        }
      }
      if (l < cw) b[c0 + l] = bi;
    } else if (k + 1 < nt) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int r = t + h * NW;
        if (r < c0) {
          double a0 = 0.0, a1 = 0.0;
#pragma unroll
          for (int j = 0; j < 16; j += 2) {
            if (FULL || j < cw1) a0 += v[16 * h + j] * b[c1 + j];
            if (FULL || j + 1 < cw1) a1 += v[16 * h + j + 1] * b[c1 + j + 1];
          }
          b[r] -= a0 + a1;
        }
      }
    }
    fetch(k - 2, v, rl);
    __syncthreads();
  };
  double va[32], vb[32], ra = 0.0, rb = 0.0;
  fetch(nt - 1, va, ra);
  fetch(nt - 2, vb, rb);
  const std::integral_constant<bool, true> full;
  const std::integral_constant<bool, false> part;
  // blocks nt - 1 and nt - 2 in the general form, then the full ones two at a time (their register
  // sets keep alternating: block k uses va when nt - 1 - k is even)
  step(nt - 1, va, ra, part);
  if (nt - 2 >= 0) step(nt - 2, vb, rb, part);
  int k = nt - 3;
  for (; k - 1 >= 0; k -= 2) {
    step(k, va, ra, full);
    step(k - 1, vb, rb, full);
  }
  if (k >= 0) step(k, va, ra, full);
}

// the one-barrier sweeps need every row below (above) a block on at most two threads of the
// wavefronts past the first
template <int NT>
__host__ __device__ constexpr bool sweep_lds_ok(int m) {
  return NT >= 128 && m - 16 <= 2 * (NT - 64);
}

// The blocked sweeps with b staged in LDS for m <= kMfmaMaxN (the in-block eliminations and the
// row updates then read and write b without global round trips), in place in b otherwise.  PF
// selects the one-barrier sweeps with L loaded two blocks ahead (fsub_lds / bsub_t_lds): their
// 248 VGPRs would cap every kernel that links them at one wavefront per SIMD, so only kernels that
// run at that occupancy anyway (the elastic kernels: one 256-thread workgroup per CU) take them.
template <int NT, bool PF = false>
__device__ void block_fsub_tiled(const double* L, const double* rcp, int m, double* b) {
  __shared__ double sb[kMfmaMaxN];
  if (m > kMfmaMaxN) {
    fsub_blocks<NT>(L, rcp, m, b);
    return;
  }
  for (int i = threadIdx.x; i < m; i += NT) sb[i] = b[i];
  __syncthreads();
  if (PF && sweep_lds_ok<NT>(m)) fsub_lds<NT>(L, rcp, m, sb);
  else fsub_blocks<NT>(L, rcp, m, sb);
  for (int i = threadIdx.x; i < m; i += NT) b[i] = sb[i];
  __syncthreads();
}

template <int NT, bool PF = false>
__device__ void block_bsub_t_tiled(const double* L, const double* rcp, int m, double* b) {
  __shared__ double sb[kMfmaMaxN];
  if (m > kMfmaMaxN) {
    bsub_t_blocks<NT>(L, rcp, m, b);
    return;
  }
  for (int i = threadIdx.x; i < m; i += NT) sb[i] = b[i];
  __syncthreads();
  if (PF && sweep_lds_ok<NT>(m)) bsub_t_lds<NT>(L, rcp, m, sb);
  else bsub_t_blocks<NT>(L, rcp, m, sb);
  for (int i = threadIdx.x; i < m; i += NT) b[i] = sb[i];
  __syncthreads();
}

// W = L^{-1} A^T (n x me, row-major) by blocked forward substitution with me right-hand sides:
// wavefront 0 eliminates each 16-row block (one lane per column of W), then every wavefront
// applies the block to its share of the rows below.
template <int NT>
__device__ __forceinline__ void block_w_tiled(const double* L, const double* rcp, int n, int me, double* Wm) {
  constexpr int W = NT / 64;
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  const int nt = (n + 15) / 16;
  for (int k = 0; k < nt; ++k) {
    const int c0 = 16 * k, cw = n - c0 < 16 ? n - c0 : 16;
    if (w == 0) {
      for (int cg = 0; cg < me; cg += 64) {
        const int c = cg + l;
        double wj[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          double wi = (i < cw && c < me) ? Wm[(c0 + i) * me + c] : 0.0;
#pragma unroll
          for (int j = 0; j < i; ++j) wi -= (i < cw ? L[(c0 + i) * n + c0 + j] : 0.0) * wj[j];
          wi *= i < cw ? rcp[c0 + i] : 0.0;
          wj[i] = wi;
          if (i < cw && c < me) Wm[(c0 + i) * me + c] = wi;
        }
      }
    }
    __syncthreads();
    for (int cg = 0; cg < me; cg += 64) {
      const int c = cg + l;
      if (c < me) {
        for (int r = c0 + 16 + w; r < n; r += W) {
          const double* lr = L + r * n + c0;
          double a0 = 0.0, a1 = 0.0;
          int j = 0;
          for (; j + 2 <= cw; j += 2) {
            a0 += lr[j] * Wm[(c0 + j) * me + c];
            a1 += lr[j + 1] * Wm[(c0 + j + 1) * me + c];
          }
          if (j < cw) a0 += lr[j] * Wm[(c0 + j) * me + c];
          Wm[r * me + c] -= a0 + a1;
        }
      }
    }
    __syncthreads();
  }
}

// W = L^{-1} A^T with W (n x me, row-major) in LDS, L from block_cholesky_tiled.  Wavefront 0
// eliminates each 16-row block (one lane per column of W, as block_w_tiled); the rows below then
// receive -L[rows, block] W[block, :] as 16x16 tiles, four v_mfma_f64_16x16x4f64 each, spread over
// the wavefronts: one batch of L loads per tile instead of one dependent global round trip per row.
template <int NT>
__device__ void block_w_lds(const double* L, const double* rcp, int n, int me, double* Wl) {
  constexpr int W = NT / 64;
  const int tid = threadIdx.x, w = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63, rq = l >> 4, cq = l & 15;
  const int nt = (n + 15) / 16, ct = (me + 15) / 16;
  const QPX_GLOBAL_AS double* Lg = as_global(L);
  for (int k = 0; k < nt; ++k) {
    const int c0 = 16 * k, cw = n - c0 < 16 ? n - c0 : 16;
    if (w == 0) {
      // the diagonal block spread over the wavefront: lane cq + 16 t4 holds L[c0 + cq, c0 + 4 t4 + u]
      // in lb[u]; L[c0 + i, c0 + j] is then lane i + 16 (j / 4)'s lb[j % 4] (clamped indices)
      const int ri = c0 + (cq < cw ? cq : cw - 1);
      double lb[4], rc[16];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int j = 4 * rq + u;
        lb[u] = Lg[ri * n + c0 + (j < cw ? j : cw - 1)];
      }
#pragma unroll
      for (int i = 0; i < 16; ++i) rc[i] = as_global(rcp)[c0 + (i < cw ? i : cw - 1)];
      for (int cg = 0; cg < me; cg += 64) {
        const int c = cg + l;
        double wj[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          double wi = (i < cw && c < me) ? Wl[(c0 + i) * me + c] : 0.0;
#pragma unroll
          for (int j = 0; j < i; ++j) wi -= (i < cw ? readlane_d(lb[j & 3], i + 16 * (j >> 2)) : 0.0) * wj[j];
          wi *= i < cw ? rc[i] : 0.0;
          wj[i] = wi;
          if (i < cw && c < me) Wl[(c0 + i) * me + c] = wi;
        }
      }
    }
    __syncthreads();
    const int T = (nt - k - 1) * ct;
    for (int idx = w; idx < T; idx += W) {
      const int R = k + 1 + idx / ct, C = idx - (idx / ct) * ct;
      const int ia = 16 * R + cq, jb = 16 * C + cq;
      const int iac = ia < n ? ia : n - 1, jbc = jb < me ? jb : me - 1;
      double av[4], bv[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int kk = 4 * s + rq;  // column of the block
        const int kc = kk < cw ? kk : cw - 1;
        const double lv = Lg[iac * n + c0 + kc], wv = Wl[(c0 + kc) * me + jbc];
        av[s] = (ia < n && kk < cw) ? -lv : 0.0;
        bv[s] = (kk < cw && jb < me) ? wv : 0.0;
      }
      d4v acc;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 16 * R + 4 * r + rq;
        const double wv = Wl[(i < n ? i : n - 1) * me + jbc];
        acc[r] = (i < n && jb < me) ? wv : 0.0;
      }
#pragma unroll
      for (int s = 0; s < 4; ++s) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s], bv[s], acc, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 16 * R + 4 * r + rq;
        if (i < n && jb < me) Wl[i * me + jb] = acc[r];
      }
    }
    __syncthreads();
  }
}

// Lower triangle of F = W^T W (me x me, row-major, into global F) from W in LDS: 16x16 tiles
// (C >= D) accumulated over the n rows of W with v_mfma_f64_16x16x4f64, one tile per wavefront
// at a time.
template <int NT>
__device__ void wtw_mfma(const double* Wl, int n, int me, double* F) {
  constexpr int W = NT / 64;
  const int tid = threadIdx.x, w = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63, rq = l >> 4, cq = l & 15;
  const int ct = (me + 15) / 16, T = ct * (ct + 1) / 2;
  for (int idx = w; idx < T; idx += W) {
    int C = 0;
    while ((C + 1) * (C + 2) / 2 <= idx) ++C;
    const int D = idx - C * (C + 1) / 2;
    const int ca = 16 * C + cq, cb = 16 * D + cq;
    d4v acc = d4v{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < n; k0 += 4) {
      const int k = k0 + rq;
      const double av = (k < n && ca < me) ? Wl[k * me + ca] : 0.0;
      const double bv = (k < n && cb < me) ? Wl[k * me + cb] : 0.0;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = 16 * C + 4 * r + rq;
      if (i < me && cb <= i) F[i * me + cb] = acc[r];
    }
  }
}

// Triangular solves with a factor from block_cholesky[_tiled]: blocked from m = 64 on.
template <int NT, bool PF = false>
__device__ void block_fsub(const double* L, const double* rcp, int m, double* b) {
  if (m >= kMfmaMinN) block_fsub_tiled<NT, PF>(L, rcp, m, b);
  else block_fsub_scalar<NT>(L, rcp, m, b);
}

template <int NT, bool PF = false>
__device__ void block_bsub_t(const double* L, const double* rcp, int m, double* b) {
  if (m >= kMfmaMinN) block_bsub_t_tiled<NT, PF>(L, rcp, m, b);
  else block_bsub_t_scalar<NT>(L, rcp, m, b);
}

template <int NT, typename GT>
__device__ __forceinline__ void stage_factor_g(const View& v, int dmode, const GT* G) {
  const int n = v.n, me = v.me, mi = v.mi, tid = threadIdx.x;
  if (dmode != kDgiven)
    for (int k = tid; k < mi; k += NT) v.D[k] = dmode == kDunit ? 1.0 : v.z[k] / v.s[k];
  __syncthreads();
  const bool tiled = n >= kMfmaMinN && n <= kMfmaMaxN;
  const bool fused_w = me > 0 && w_fused(n, me, NT);
  {
    if (tiled) {
      QPX_DYN_LDS(qpx_smem);
      if constexpr (NT == 256) gtdg_sb<GT>(v.Q, v.D, v.H, n, mi, reinterpret_cast<double*>(qpx_smem), G);
      else gtdg_mfma<NT>(v.Q, v.D, v.H, n, mi, reinterpret_cast<double*>(qpx_smem), G);
    } else {
      for (int e = tid; e < n * n; e += NT) {
        const int i = e / n, j = e - i * n;
        if (j > i) continue;
        double acc = 0.0;
        for (int k = 0; k < mi; ++k) acc += (double)G[k * n + i] * (v.D[k] * (double)G[k * n + j]);
        v.H[e] = v.Q[e] + acc;
      }
    }
    __syncthreads();
    if (tiled) {
      QPX_DYN_LDS(qpx_smem);
      double* lds = reinterpret_cast<double*>(qpx_smem);
      double* Wl = fused_w ? lds + 2 * panel_words(n) : nullptr;
      // W starts as A^T beside the panels (past the finished G^T D G chunk)
      if (fused_w)
        for (int e = tid; e < n * me; e += NT) {
          const int i = e / me, c = e - i * me;
          Wl[e] = v.A[c * n + i];
        }
      __syncthreads();
      block_cholesky_tiled<NT>(v.H, n, v.hrcp, lds, Wl, me,
                               lds + 2 * panel_words(n) + (fused_w ? n * me : 0));
    } else {
      block_cholesky<NT>(v.H, n, v.hdiag, v.hrcp);
    }
  }
  if (me == 0) return;
  if (fused_w) {
    // F = W^T W on the matrix cores; W is copied to the workspace for the solves
    QPX_DYN_LDS(qpx_smem);
    const double* Wl = reinterpret_cast<const double*>(qpx_smem) + 2 * panel_words(n);
    wtw_mfma<NT>(Wl, n, me, v.F);
    for (int e = tid; e < n * me; e += NT) v.W[e] = Wl[e];
    __syncthreads();
    // the Cholesky panels' LDS is free now
    if (me <= kSmallChol) block_cholesky_small<NT>(v.F, me, v.fdiag, v.frcp, reinterpret_cast<double*>(qpx_smem));
    else block_cholesky<NT>(v.F, me, v.fdiag, v.frcp);
    return;
  }
  if (w_in_lds(n, me, NT)) {
    // W in LDS (aliasing the finished G^T D G chunk and Cholesky panel), its trailing updates and
    // F = W^T W on the matrix cores; W is copied to the workspace for the solves
    QPX_DYN_LDS(qpx_smem);
    double* Wl = reinterpret_cast<double*>(qpx_smem);
    __syncthreads();
    for (int e = tid; e < n * me; e += NT) {
      const int i = e / me, c = e - i * me;
      Wl[e] = v.A[c * n + i];
    }
    __syncthreads();
    block_w_lds<NT>(v.H, v.hrcp, n, me, Wl);
    wtw_mfma<NT>(Wl, n, me, v.F);
    for (int e = tid; e < n * me; e += NT) v.W[e] = Wl[e];
    __syncthreads();
    // W's LDS copy is no longer read
    if (me <= kSmallChol) block_cholesky_small<NT>(v.F, me, v.fdiag, v.frcp, Wl);
    else block_cholesky<NT>(v.F, me, v.fdiag, v.frcp);
    return;
  }
  // W = L^{-1} A^T, all me columns at once: column-oriented forward substitution.
  for (int e = tid; e < n * me; e += NT) {
    const int i = e / me, c = e - i * me;
    v.W[e] = v.A[c * n + i];
  }
  __syncthreads();
  if (tiled) block_w_tiled<NT>(v.H, v.hrcp, n, me, v.W);
  for (int j = 0; j < (tiled ? 0 : n); ++j) {
    // row j of W is final once rows < j have been eliminated; scale it after its last use
    const double rj = v.hrcp[j];
    for (int e = tid; e < (n - j - 1) * me; e += NT) {
      const int a = e / me, c = e - a * me, i = j + 1 + a;
      v.W[i * me + c] -= v.H[i * n + j] * (v.W[j * me + c] * rj);
    }
    __syncthreads();
    for (int c = tid; c < me; c += NT) v.W[j * me + c] *= rj;
    __syncthreads();
  }
  for (int e = tid; e < me * me; e += NT) {
    const int c = e / me, d = e - c * me;
    if (d > c) continue;
    double acc = 0.0;
    for (int i = 0; i < n; ++i) acc += v.W[i * me + c] * v.W[i * me + d];
    v.F[e] = acc;
  }
  __syncthreads();
  block_cholesky<NT>(v.F, me, v.fdiag, v.frcp);
}

template <int NT>
__device__ __forceinline__ void stage_factor(const View& v, int dmode) {
  stage_factor_g<NT, double>(v, dmode, v.G);
}

// Core of the reduced solve: given p1 (in v.u, length n) and v4, produce dy (v.dy) and dx.
// dy = F^{-1}(A H^{-1} p1 - v4), dx = H^{-1}(p1 - A^T dy) with A H^{-1} = W^T L^{-1}.
template <int NT, bool PF = false>
__device__ void reduced_solve(const View& v, const double* v4, double* dx) {
  const int n = v.n, me = v.me, tid = threadIdx.x;
  block_fsub<NT, PF>(v.H, v.hrcp, n, v.u);  // u = L^{-1} p1
  if (me > 0) {
    if (n >= kMfmaMinN && n <= kMfmaMaxN && me <= 64) {  // t = W^T u - v4 as column sums
      QPX_DYN_LDS(qpx_smem);
      double acc[1] = {0.0};
      colsum_acc<NT, 1>(v.W, n, me, v.u, acc);
      colsum_finish<NT, 1>(acc, me, v4, -1.0, 1.0, v.t, reinterpret_cast<double*>(qpx_smem));
    } else {
      for (int c = tid; c < me; c += NT) {
        double acc = 0.0;
        for (int i = 0; i < n; ++i) acc += v.W[i * me + c] * v.u[i];
        v.t[c] = acc - v4[c];
      }
      __syncthreads();
    }
    block_fsub<NT>(v.F, v.frcp, me, v.t);
    block_bsub_t<NT>(v.F, v.frcp, me, v.t);
    for (int c = tid; c < me; c += NT) v.dy[c] = v.t[c];
    for (int i = tid; i < n; i += NT) {
      double acc = 0.0;
      for (int c = 0; c < me; ++c) acc += v.W[i * me + c] * v.t[c];
      v.u[i] -= acc;
    }
    __syncthreads();
  }
  block_bsub_t<NT, PF>(v.H, v.hrcp, n, v.u);
  for (int i = tid; i < n; i += NT) dx[i] = v.u[i];
  __syncthreads();
}

// --------------------------------------------------------------------------------------------
// Stage: solve the KKT system for right-hand side (v1, v2, v3, v4) (qpax/pdip.py:49-63).
// The right-hand sides are passed negated-or-not through sign flags so callers can hand over the
// residual vectors directly: v1 = sg1*a1, v2 = sg2*a2 (+ c2 if non-null), v3 = sg3*a3,
// v4 = sg4*a4.  Writes dx, ds, dz (into the given arrays) and dy (v.dy).
// --------------------------------------------------------------------------------------------
template <int NT>
__device__ void stage_solve(const View& v, const double* a1, const double* a2, const double* a3,
                            const double* a4, double sg, double* dx, double* ds, double* dz) {
  const int n = v.n, me = v.me, mi = v.mi, tid = threadIdx.x;
  // r2' = v3 - v2/z, stored into ds temporarily; p1 = v1 + G^T (D r2')
  for (int k = tid; k < mi; k += NT) ds[k] = sg * a3[k] - (sg * a2[k]) / v.z[k];
  __syncthreads();
  for (int i = tid; i < n; i += NT) {
    double acc = 0.0;
    for (int k = 0; k < mi; ++k) acc += v.G[k * n + i] * (v.D[k] * ds[k]);
    v.u[i] = sg * a1[i] + acc;
  }
  // v4 = sg * a4 -> handled by scaling t; write it into dy scratch first
  for (int c = tid; c < me; c += NT) v.dy[c] = sg * a4[c];
  __syncthreads();
  reduced_solve<NT>(v, v.dy, dx);
  const bool wide = n >= kMfmaMinN;
  if (wide) {  // G dx into dz first
    rows_dot16<NT>(v.G, mi, n, dx, dz);
    __syncthreads();
  }
  for (int k = tid; k < mi; k += NT) {
    double acc = 0.0;
    if (wide) acc = dz[k];
    else
      for (int j = 0; j < n; ++j) acc += v.G[k * n + j] * dx[j];
    const double dsk = sg * a3[k] - acc;
    ds[k] = dsk;
    dz[k] = (sg * a2[k] - v.z[k] * dsk) / v.s[k];
  }
  __syncthreads();
}

// --------------------------------------------------------------------------------------------
// Stage: residuals of the (relaxed) KKT system and the entry convergence flag
// (qpax/pdip.py:93-100; pdip_relaxed.py:16-23 with kappa).  Returns the flag.  With `norms`,
// thread 0 also writes the infinity norms of r1, r2, r3 and r4 there (0 for an empty block, as
// the reference's debug table prints, pdip.py:286-298).
// --------------------------------------------------------------------------------------------
template <int NT>
__device__ int stage_residuals(const View& v, double kappa, double tol, double* red, double* norms = nullptr) {
  const int n = v.n, me = v.me, mi = v.mi, tid = threadIdx.x;
  const bool wide = n >= kMfmaMinN;
  if (wide) {  // G x into r3 and A x into r4 first; Q x by columns (Q is symmetric), coalesced
    rows_dot16<NT>(v.G, mi, n, v.x, v.r3);
    rows_dot16<NT>(v.A, me, n, v.x, v.r4);
    __syncthreads();
  }
  double m1 = 0.0, m2 = 0.0, m3 = 0.0, m4 = 0.0;
  for (int i = tid; i < n; i += NT) {
    double acc = 0.0;
    if (wide)
      for (int j = 0; j < n; ++j) acc += v.Q[j * n + i] * v.x[j];
    else
      for (int j = 0; j < n; ++j) acc += v.Q[i * n + j] * v.x[j];
    acc += v.q[i];
    for (int c = 0; c < me; ++c) acc += v.A[c * n + i] * v.y[c];
    for (int k = 0; k < mi; ++k) acc += v.G[k * n + i] * v.z[k];
    v.r1[i] = acc;
    m1 = nmax(m1, fabs(acc));
  }
  for (int k = tid; k < mi; k += NT) {
    double acc = 0.0;
    if (wide) acc = v.r3[k];
    else
      for (int j = 0; j < n; ++j) acc += v.G[k * n + j] * v.x[j];
    const double r3 = acc + v.s[k] - v.h[k];
    const double r2 = v.s[k] * v.z[k] - kappa;
    v.r2[k] = r2;
    v.r3[k] = r3;
    m2 = nmax(m2, fabs(r2));
    m3 = nmax(m3, fabs(r3));
  }
  for (int c = tid; c < me; c += NT) {
    double acc = 0.0;
    if (wide) acc = v.r4[c];
    else
      for (int j = 0; j < n; ++j) acc += v.A[c * n + j] * v.x[j];
    const double r4 = acc - v.b[c];
    v.r4[c] = r4;
    m4 = nmax(m4, fabs(r4));
  }
  double m;
  if (norms) {
    m1 = block_reduce<NT>(m1, red, OpMax());
    __syncthreads();
    m2 = block_reduce<NT>(m2, red, OpMax());
    __syncthreads();
    m3 = block_reduce<NT>(m3, red, OpMax());
    __syncthreads();
    m4 = block_reduce<NT>(m4, red, OpMax());
    if (tid == 0) { norms[0] = m1; norms[1] = m2; norms[2] = m3; norms[3] = m4; }
    m = nmax(nmax(m1, m2), nmax(m3, m4));
  } else {
    m = block_reduce<NT>(nmax(nmax(m1, m2), nmax(m3, m4)), red, OpMax());
  }
  __syncthreads();
  return (m < tol) ? 1 : 0;
}

// min(1, min_i{-x_i/dx_i : dx_i < 0}) over the block (qpax/pdip.py:66-73).
template <int NT>
__device__ double block_ort(const double* x, const double* dx, int m, double* red) {
  double a = 1.0;
  for (int k = threadIdx.x; k < m; k += NT) {
    const double d = dx[k];
    if (d < 0) a = nmin(a, -x[k] / d);
  }
  return block_reduce<NT>(a, red, OpMin());
}

// Stage: Mehrotra centering (qpax/pdip.py:76-84, :107-108).  Updates r2 in place:
// r2 <- r2 - (sigma mu - ds_a dz_a).
template <int NT>
__device__ void stage_center(const View& v, double* red) {
  const int mi = v.mi, tid = threadIdx.x;
  double sz = 0.0;
  for (int k = tid; k < mi; k += NT) sz += v.s[k] * v.z[k];
  sz = block_reduce<NT>(sz, red, OpSum());
  const double mu = sz / mi;
  const double a = nmin(block_ort<NT>(v.s, v.dsa, mi, red), block_ort<NT>(v.z, v.dza, mi, red));
  double g = 0.0;
  for (int k = tid; k < mi; k += NT) g += (v.s[k] + a * v.dsa[k]) * (v.z[k] + a * v.dza[k]);
  g = block_reduce<NT>(g, red, OpSum());
  const double ratio = g / sz;
  const double sm = ratio * ratio * ratio * mu;
  for (int k = tid; k < mi; k += NT) v.r2[k] = v.r2[k] - (sm - v.dsa[k] * v.dza[k]);
  __syncthreads();
}

// Stage: line search and update (qpax/pdip.py:112-117 for the predictor-corrector step,
// pdip_relaxed.py:30-35 for the relaxed Newton step).
template <int NT>
__device__ double stage_step(const View& v, bool relax_rule, double* red) {
  const int n = v.n, me = v.me, mi = v.mi, tid = threadIdx.x;
  const double as = block_ort<NT>(v.s, v.ds, mi, red);
  const double az = block_ort<NT>(v.z, v.dz, mi, red);
  const double alpha = relax_rule ? 0.99 * nmin(nmin(1.0, 0.99 * as), 0.99 * az) : 0.99 * nmin(as, az);
  for (int i = tid; i < n; i += NT) v.x[i] += alpha * v.dx[i];
  for (int k = tid; k < mi; k += NT) { v.s[k] += alpha * v.ds[k]; v.z[k] += alpha * v.dz[k]; }
  for (int c = tid; c < me; c += NT) v.y[c] += alpha * v.dy[c];
  __syncthreads();
  return alpha;
}

// Stage: initialisation (qpax/pdip.py:15-35).  Factor with D = 1, solve the reduced system with
// p1 = -q + G^T h and v4 = b, then shift s and z into the positive orthant.
template <int NT>
__device__ void stage_init(const View& v, double* red) {
  const int n = v.n, me = v.me, mi = v.mi, tid = threadIdx.x;
  stage_factor<NT>(v, kDunit);
  for (int i = tid; i < n; i += NT) {
    double acc = -v.q[i];
    for (int k = 0; k < mi; ++k) acc += v.G[k * n + i] * v.h[k];
    v.u[i] = acc;
  }
  __syncthreads();
  reduced_solve<NT>(v, v.b, v.x);
  for (int c = tid; c < me; c += NT) v.y[c] = v.dy[c];
  double zmax = -INFINITY, zmin = INFINITY;
  for (int k = tid; k < mi; k += NT) {
    double acc = 0.0;
    for (int j = 0; j < n; ++j) acc += v.G[k * n + j] * v.x[j];
    const double zk = acc - v.h[k];
    v.z[k] = zk;
    zmax = nmax(zmax, zk);
    zmin = nmin(zmin, zk);
  }
  zmax = block_reduce<NT>(zmax, red, OpMax());
  zmin = block_reduce<NT>(zmin, red, OpMin());
  const double alpha_p = zmax, alpha_d = -zmin;
  for (int k = tid; k < mi; k += NT) {
    const double zk = v.z[k];
    v.s[k] = (alpha_p < 0) ? -zk : -zk + (1 + alpha_p);
    v.z[k] = (alpha_d >= 0) ? zk + (1 + alpha_d) : zk;
  }
  __syncthreads();
}

// One predictor-corrector iteration (qpax/pdip.py:87-119) on an active problem.  With `trace`,
// thread 0 writes the entry residual norms (r1, r2, r3, r4) and the step length there;
// `debug_step` takes the step length rule of the reference's debug loop,
// 0.99 min(1, 0.99 ort(s), 0.99 ort(z)) (pdip.py:281), instead of 0.99 min(ort(s), ort(z)).
template <int NT>
__device__ void pc_iteration(const View& v, double tol, double* red, double* trace = nullptr, bool debug_step = false) {
  const int conv = stage_residuals<NT>(v, 0.0, tol, red, trace);
  stage_factor<NT>(v, kDzs);
  stage_solve<NT>(v, v.r1, v.r2, v.r3, v.r4, -1.0, v.dx, v.dsa, v.dza);
  stage_center<NT>(v, red);
  stage_solve<NT>(v, v.r1, v.r2, v.r3, v.r4, -1.0, v.dx, v.ds, v.dz);
  const double alpha = stage_step<NT>(v, debug_step, red);
  if (trace && threadIdx.x == 0) trace[4] = alpha;
  if (threadIdx.x == 0) { v.ctl[kConv] = conv; v.ctl[kIter] += 1; }
  __syncthreads();
}

// One relaxed Newton iteration (qpax/pdip_relaxed.py:7-52).
template <int NT>
__device__ void relax_iteration(const View& v, double tol, double kappa, double* red) {
  const int conv = stage_residuals<NT>(v, kappa, tol, red);
  stage_factor<NT>(v, kDzs);
  stage_solve<NT>(v, v.r1, v.r2, v.r3, v.r4, -1.0, v.dx, v.ds, v.dz);
  stage_step<NT>(v, true, red);
  if (threadIdx.x == 0) { v.ctl[kRelaxConv] = conv; v.ctl[kRelaxIter] += 1; }
  __syncthreads();
}

}  // namespace qpx
