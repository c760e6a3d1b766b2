// Fused solver for mid-size dense QPs: one workgroup of W wavefronts owns one QP.
//
// The augmented system is the one of the wavefront kernel (qpx_wave.hpp): rows 0..n-1 are x,
// n..np-1 identity padding (np = n rounded up to a multiple of 4), np..m-1 the equality
// multipliers y; K = [[Q + G^T diag(z/s) G, A^T], [A, 0]] is factored as one quasi-definite
// L D L^T in panels of 4 pivots (the cvxopt reduction of qpax/pdip.py:38-63 as one elimination).
// Thread t owns row t of the augmented vectors and rows t, t + 64W, ... of the m_ineq-vectors.
//
// The lower triangle of K is held as 16x16 float64 matrix-core tiles spread over the wavefronts:
// tile (I, J), J <= I, with row-major index I(I+1)/2 + J, lives in slot s of wavefront w when the
// index is w + W s, which balances the tiles to within one per wavefront.  Per factorisation:
//   * the tiles start from K's constant part, kept in the problem's global workspace;
//   * G^T D G accumulates with v_mfma_f64_16x16x4f64 while G streams through LDS in chunks of 16
//     rows (double-buffered, one barrier per chunk); G is read once from the workspace copy
//     (rows with infinite h zeroed, columns padded with zeros to 16 * ceil(m / 16));
//   * every panel of 4 pivots is extracted to LDS, eliminated redundantly by every thread from
//     the 4 x 4 diagonal block (so no thread waits on another for its multipliers), and applied
//     to the trailing tiles as one rank-4 MFMA per tile;
//   * the unit-lower factor is written packed to LDS (row i at i(i+1)/2) for the two triangular
//     sweeps, which run wavefront by wavefront: the owner of a 64-row block eliminates inside it
//     with lane broadcasts, then every later (earlier, for L^T) wavefront applies the block.
// G x and G^T v read the workspace copy of G directly: 16 lanes per row of G (coalesced, then a
// DPP reduction) for G x, one thread per column for G^T v.
#pragma once
#include "qpx_args.hpp"
#include "qpx_core.hpp"
#include "qpx_wave.hpp"

namespace qpx {
namespace group {

using wave::d4;
using wave::recip;
using wave::rl;

constexpr int KC = 32;   // rows of G per LDS chunk
constexpr int LDB = 17;  // row stride (words) of the multiplier buffer: conflict-free operand reads

__host__ __device__ constexpr int mic_of(int W) { return 512 / (64 * W) > 1 ? 512 / (64 * W) : 1; }
__host__ __device__ constexpr int smax_of(int NTI, int W) { return (NTI * (NTI + 1) / 2 + W - 1) / W; }
__host__ __device__ inline int ntiles(int m) { return (m + 15) / 16; }
// Row stride of G (workspace copy and LDS chunks): 16 * nti rounded to an odd multiple of 16, so
// that the four rows read by one matrix-core operand load fall in disjoint LDS banks.
__host__ __device__ inline int g_stride(int m) { return 16 * (ntiles(m) | 1); }

struct GLds {
  int off_L, off_A, off_B, off_pv, off_vb, off_vc, off_vz, off_d0, off_red, off_part, bytes;
};

// LDS carve-up for one problem (float64 words): the packed factor, which also holds the two G
// chunks (storage type) during G^T D G; the panel extraction buffer, aliased by the augmented
// vectors of the solves (x and dx in the backward); the multipliers of one tile column and its
// pivots; two m_ineq-vectors; the reference diagonal; reduction scratch; partial sums of the
// column products.
__host__ __device__ inline GLds group_lds(int m, int mi, int NT, int ts) {
  const int rows = 16 * ntiles(m), ldg = g_stride(m);
  const int fac = m * (m + 1) / 2;
  const int stage = (2 * KC * ldg * ts + 7) / 8;
  const int vec = 2 * (NT > ldg ? NT : ldg);
  GLds w;
  int o = 0;
  w.off_L = o; o += (fac > stage ? fac : stage);
  o = (o + 1) & ~1;
  w.off_A = o; o += (4 * rows > vec ? 4 * rows : vec);
  w.off_B = o; o += LDB * rows + 1;
  o = (o + 1) & ~1;
  w.off_pv = o; o += 16;
  w.off_vb = o; o += (mi + 1) & ~1;
  w.off_vc = o; o += (mi + 1) & ~1;
  w.off_vz = o; o += (mi + KC - 1) / KC * KC;
  w.off_d0 = o; o += rows;
  w.off_red = o; o += 16;
  w.off_part = o; o += NT;
  w.bytes = 8 * o;
  return w;
}

// Global workspace words of one problem: K's constant part (m x m) and the padded copy of G
// (m_ineq x 16 ceil(m/16), counted as float64 words whatever the storage type).
__host__ __device__ inline int64_t group_words(int n, int me, int mi) {
  const int64_t m = wave::padded_n(n) + me, ldg = g_stride((int)m);
  return round_up(m * m, 64) + round_up((int64_t)mi * ldg, 64);
}

template <int MIC>
struct GSt {
  double qa, xa, ra, da, rdpiv;
  double h[MIC];
  bool hinf[MIC];
  double s[MIC], z[MIC], D[MIC], r2[MIC], r3[MIC], dsa[MIC], dza[MIC], ds[MIC], dz[MIC];
};

// Tile slots of the calling wavefront.  Tiles are numbered column-major over the lower triangle
// of an NTI x NTI tile grid (column J holds rows J..NTI-1); slot s holds tile w + W s.  A slot is
// valid when that tile exists and lies inside this problem's nti x nti grid.
template <typename TS, int NTI, int W>
struct GCtx {
  static constexpr int SMAX = smax_of(NTI, W);
  int n, np, me, mi, m, nti, ldg;
  int t, w, l, rq, cq;
  int sI[SMAX], sJ[SMAX];
  bool sv[SMAX];
  double* KQ;   // workspace: K's constant part, entry (i, j) at KQ[j * m + i]
  TS* Gw;       // workspace: G with rows of infinite h zeroed, row stride ldg, zero columns >= n
  double* Lp;   // LDS
  TS* Gst;      // LDS, aliases Lp
  double* pbA;  // aliased by col
  double* col;
  double* pbB;
  double* pv;
  double* vb;
  double* vc;
  double* vz;   // z (zero-padded to whole G chunks) for the fused residual products
  double* d0;
  double* red;
  double* part;
};

template <int W, class Op>
__device__ __forceinline__ double breduce(double v, Op op, double* red) {
  v = wave::wreduce(v, op);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = red[0];
#pragma unroll
  for (int i = 1; i < W; ++i) r = op(r, red[i]);
  return r;
}

// Sum over the 16 lanes of each DPP row; every lane of the row receives it.
__device__ __forceinline__ double row16_sum(double v) {
  v += wave::dpp<0xB1>(v);
  v += wave::dpp<0x4E>(v);
  v += wave::dpp<0x141>(v);
  v += wave::dpp<0x140>(v);
  return v;
}

template <typename TS, int NTI, int W>
__device__ __forceinline__ GCtx<TS, NTI, W> make_gctx(int n, int me, int mi, double* ws) {
  QPX_DYN_LDS(qpx_smem);
  constexpr int NT = 64 * W, TT = NTI * (NTI + 1) / 2, SMAX = smax_of(NTI, W);
  GCtx<TS, NTI, W> c;
  c.n = n; c.np = wave::padded_n(n); c.me = me; c.mi = mi; c.m = c.np + me;
  c.nti = ntiles(c.m); c.ldg = g_stride(c.m);
  c.t = threadIdx.x;
  c.w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  c.l = threadIdx.x & 63; c.rq = c.l >> 4; c.cq = c.l & 15;
#pragma unroll
  for (int s = 0; s < SMAX; ++s) {
    const int idx = c.w + W * s;
    int J = 0, start = 0;
    while (J < NTI - 1 && start + (NTI - J) <= idx) { start += NTI - J; ++J; }
    const int I = J + (idx - start);
    const bool ok = idx < TT && I < c.nti;
    c.sv[s] = ok;
    c.sI[s] = ok ? I : 0;
    c.sJ[s] = ok ? J : 0;
  }
  const GLds L = group_lds(c.m, mi, NT, (int)sizeof(TS));
  double* base = reinterpret_cast<double*>(qpx_smem);
  c.Lp = base + L.off_L; c.Gst = reinterpret_cast<TS*>(c.Lp);
  c.pbA = base + L.off_A; c.col = c.pbA;
  c.pbB = base + L.off_B; c.pv = base + L.off_pv;
  c.vb = base + L.off_vb; c.vc = base + L.off_vc; c.vz = base + L.off_vz;
  c.d0 = base + L.off_d0; c.red = base + L.off_red; c.part = base + L.off_part;
  c.KQ = ws;
  c.Gw = reinterpret_cast<TS*>(ws + round_up((int64_t)c.m * c.m, 64));
  return c;
}

// Copy of the context whose sizes, thread indices and tile-slot tables the compiler cannot hoist
// out of the iteration loops (each per-iteration stage takes one): lane masks and LDS addresses
// derived from them are recomputed per stage instead of being kept live across the whole solve.
template <typename TS, int NTI, int W>
__device__ __forceinline__ GCtx<TS, NTI, W> fresh(const GCtx<TS, NTI, W>& c) {
  constexpr int SMAX = smax_of(NTI, W);
  GCtx<TS, NTI, W> o = c;
  o.n = opaque_s(c.n); o.np = opaque_s(c.np); o.me = opaque_s(c.me);
  o.mi = opaque_s(c.mi); o.m = opaque_s(c.m); o.nti = opaque_s(c.nti);
  o.ldg = opaque_s(c.ldg); o.w = opaque_s(c.w);
  o.t = opaque_v(c.t); o.l = opaque_v(c.l); o.rq = opaque_v(c.rq); o.cq = opaque_v(c.cq);
#pragma unroll
  for (int s = 0; s < SMAX; ++s) {
    o.sI[s] = opaque_s(c.sI[s]);
    o.sJ[s] = opaque_s(c.sJ[s]);
    o.sv[s] = opaque_s(c.sv[s] ? 1 : 0) != 0;
  }
  return o;
}

// Load and sanitise (qpax/pdip.py:166-178, :223): the workspace copy of G, K's constant part
// [[Q_sym, A^T], [A, 0]] with identity padding rows, h, q and -b.
template <typename TS, int NTI, int W>
__device__ __forceinline__ void load(const GCtx<TS, NTI, W>& c, GSt<mic_of(W)>& st, const TS* Q, const TS* q, const TS* A,
                     const TS* b, const TS* G, const TS* h) {
  constexpr int NT = 64 * W, MIC = mic_of(W);
  const int n = c.n, np = c.np, mi = c.mi, m = c.m, t = c.t, ldg = c.ldg;
  // the padded workspace copy of G, 8 rows per trip with their loads in flight
  for (int k0 = 0; k0 < mi; k0 += 8) {
    for (int j = t; j < ldg; j += NT) {
      TS v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int k = k0 + u;
        v[u] = (k < mi && j < n) ? G[(int64_t)k * n + j] : (TS)0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int k = k0 + u;
        if (k < mi) {
          const bool inf = isinf((double)h[k]);
          c.Gw[(int64_t)k * ldg + j] = inf ? (TS)0 * v[u] : v[u];  // diag(1 * ~mask) @ G (pdip.py:176)
        }
      }
    }
  }
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = NT * cc + t;
    const double hk = k < mi ? (double)h[k] : 0.0;
    st.hinf[cc] = isinf(hk);
    st.h[cc] = st.hinf[cc] ? 0.0 : hk;
    st.s[cc] = 1.0;
    st.z[cc] = 1.0;
  }
  if (t < m) {  // K's constant part, 8 columns per trip
    for (int j0 = 0; j0 < m; j0 += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int j = j0 + u;
        double val = 0.0;
        if (j < m) {
          if (t < n) {
            if (j < n) val = 0.5 * ((double)Q[t * n + j] + (double)Q[j * n + t]);
            else if (j >= np) val = (double)A[(j - np) * n + t];
          } else if (t < np) {
            val = (j == t) ? 1.0 : 0.0;
          } else if (j < n) {
            val = (double)A[(t - np) * n + j];
          }
        }
        v[u] = val;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (j0 + u < m) c.KQ[(int64_t)(j0 + u) * m + t] = v[u];
    }
  }
  st.qa = t < n ? (double)q[t] : ((t >= np && t < m) ? -(double)b[t - np] : 0.0);
  st.xa = 0.0;
  __threadfence_block();
  __syncthreads();
}

// out[k] = sum_j G[k][j] * col[j] for every row k < mi (col zero beyond m); 16 lanes per row,
// two rows per lane group per trip so that twice the loads are in flight.
template <typename TS, int NTI, int W>
__device__ __forceinline__ void g_rows(const GCtx<TS, NTI, W>& cin, double* out) {
  const GCtx<TS, NTI, W> c = fresh(cin);
  const int mi = c.mi, ldg = c.ldg, nti = c.nti, rq = c.rq, cq = c.cq;
  for (int kb = 4 * c.w; kb < mi; kb += 8 * W) {
    const int k0 = kb + rq, k1 = kb + 4 * W + rq;
    const TS* g0 = c.Gw + (int64_t)(k0 < mi ? k0 : mi - 1) * ldg + cq;
    const TS* g1 = c.Gw + (int64_t)(k1 < mi ? k1 : mi - 1) * ldg + cq;
    double a0 = 0.0, a1 = 0.0;
#pragma unroll
    for (int u = 0; u < NTI; ++u) {
      if (u < nti) {
        const double xv = c.col[16 * u + cq];
        a0 += (double)g0[16 * u] * xv;
        a1 += (double)g1[16 * u] * xv;
      }
    }
    a0 = row16_sum(a0);
    a1 = row16_sum(a1);
    if (cq == 0 && k0 < mi) out[k0] = a0;
    if (cq == 0 && k1 < mi) out[k1] = a1;
  }
}

// Column products out_t = sum_r M[r][t] * v[r] over rows r < R, for every column t < C (M row-major
// with row stride ld, in global memory), returned to thread t (0 for t >= C).  Columns go to
// groups of 64 lanes, the rows are split over the NT / ceil64(C) thread groups so that every
// thread streams a share of M (coalesced across a wavefront); the partial sums meet in LDS in a
// fixed order.
template <int W, typename TM>
__device__ __forceinline__ double col_dot(const TM* M, int64_t ld, int R, int C, const double* v, double* part) {
  constexpr int NT = 64 * W;
  const int t = threadIdx.x;
  const int cp = (C + 63) & ~63;
  const int P = NT / cp > 0 ? NT / cp : 1;
  const int j = t % cp, pt = t / cp;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  if (pt < P && j < C) {
    const TM* g = M + j;
    int r = pt;
    for (; r + 3 * P < R; r += 4 * P) {
      a0 += (double)g[(int64_t)r * ld] * v[r];
      a1 += (double)g[(int64_t)(r + P) * ld] * v[r + P];
      a2 += (double)g[(int64_t)(r + 2 * P) * ld] * v[r + 2 * P];
      a3 += (double)g[(int64_t)(r + 3 * P) * ld] * v[r + 3 * P];
    }
    for (; r < R; r += P) a0 += (double)g[(int64_t)r * ld] * v[r];
  }
  __syncthreads();
  if (pt < P) part[pt * cp + j] = (a0 + a1) + (a2 + a3);
  __syncthreads();
  double out = 0.0;
  if (t < C) {
    for (int q = 0; q < P; ++q) out += part[q * cp + t];
  }
  return out;
}

template <typename TS, int NTI, int W>
__device__ __forceinline__ double g_col(const GCtx<TS, NTI, W>& cin, const double* v) {
  const GCtx<TS, NTI, W> c = fresh(cin);
  return col_dot<W>(c.Gw, c.ldg, c.mi, c.n, v, c.part);
}

// Residuals of the (relaxed) KKT system (qpax/pdip.py:93-100) in two halves around the
// factorisation, which streams G through LDS anyway and forms G x and G^T z on the way.
// Before: x to the broadcast vector, z to LDS, and r1's constant part q + K_const [x; y].
template <typename TS, int NTI, int W>
__device__ __forceinline__ void residual_pre(const GCtx<TS, NTI, W>& cin, GSt<mic_of(W)>& st) {
  const GCtx<TS, NTI, W> c = fresh(cin);
  constexpr int NT = 64 * W, MIC = mic_of(W);
  const int m = c.m, mi = c.mi, t = c.t;
  if (t < c.ldg) c.col[t] = t < m ? st.xa : 0.0;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = NT * cc + t;
    if (k < (mi + KC - 1) / KC * KC) c.vz[k] = k < mi ? st.z[cc] : 0.0;
  }
  __syncthreads();
  // K's constant part times [x; y] (KQ[j m + i] = K[i][j])
  st.ra = st.qa + col_dot<W>(c.KQ, m, m, m, c.col, c.part);
}

// After: r1 = q + K_const [x; y] + G^T z (gz: this thread's column of G^T z), r2 = s z - kappa,
// r3 = G x + s - h; returns the infinity norm (NaN-propagating).
template <typename TS, int NTI, int W>
__device__ __forceinline__ double residual_post(const GCtx<TS, NTI, W>& cin, GSt<mic_of(W)>& st, double gz,
                                                double kappa) {
  const GCtx<TS, NTI, W> c = fresh(cin);
  constexpr int NT = 64 * W, MIC = mic_of(W);
  const int m = c.m, mi = c.mi, t = c.t;
  st.ra = t < m ? st.ra + gz : 0.0;
  double nrm = fabs(st.ra);
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = NT * cc + t;
    if (k < mi) {
      st.r3[cc] = c.vc[k] + st.s[cc] - st.h[cc];
      st.r2[cc] = st.s[cc] * st.z[cc] - kappa;
      nrm = nmax(nrm, nmax(fabs(st.r2[cc]), fabs(st.r3[cc])));
    } else {
      st.r3[cc] = 0.0;
      st.r2[cc] = 0.0;
    }
  }
  return breduce<W>(nrm, OpMax(), c.red);
}

// Factorisation (qpax/pdip.py:38-46): D = z/s (or 1 for the initialisation), K = KQ + G^T D G,
// K = L D L^T, L unit lower packed in c.Lp, 1/D in st.rdpiv.  Pivot rule of the wavefront kernel:
// a pivot not beyond kPivotRel times its reference diagonal (the entry before elimination for
// the H block, the entry once all H pivots are done for the Schur block), or NaN, is replaced by
// kHugePivot with the block's sign.  Within a tile column the panels of 4 pivots update the
// column's own tiles at once; the tiles right of it receive the column's 16 pivots as one rank-16
// update (four MFMAs) once the column is done.  The trailing update uses cv = lf * pivot for
// the column values, which differs from the eliminated row values only in the last bit (and in
// already finished columns, which are never read again).
template <typename TS, int NTI, int W>
__device__ __forceinline__ void factor(const GCtx<TS, NTI, W>& cin, GSt<mic_of(W)>& st, bool unit_d,
                                       double* gz = nullptr) {
  const GCtx<TS, NTI, W> c = fresh(cin);
  constexpr int NT = 64 * W, MIC = mic_of(W), SMAX = smax_of(NTI, W);
  const int np = c.np, mi = c.mi, m = c.m, t = c.t, ldg = c.ldg, rows = 16 * c.nti, nti = c.nti;
  const int rq = c.rq, cq = c.cq;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = NT * cc + t;
    st.D[cc] = unit_d ? 1.0 : st.z[cc] / st.s[cc];
    if (k < mi) c.vb[k] = st.D[cc];
  }
  // tiles start as K's constant part
  d4 T[SMAX];
#pragma unroll
  for (int s = 0; s < SMAX; ++s) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * c.sI[s] + 4 * r + rq, cl = 16 * c.sJ[s] + cq;
      const bool in = c.sv[s] && row < m && cl < m;
      const double v = c.KQ[(int64_t)(in ? cl : 0) * m + (in ? row : 0)];
      T[s][r] = in ? v : 0.0;
    }
  }
  // G^T D G: chunks of KC rows of G staged in LDS (double-buffered); within a chunk the operands
  // of the next 4 rows load while the current MFMAs run.  Invalid slots read tile (0, 0) and
  // accumulate into their own unused registers.
  {
    constexpr int EPT = (KC * 16 * (NTI | 1) + NT - 1) / NT;
    const int csz = KC * ldg;
    const int64_t total = (int64_t)mi * ldg;
    const int nch = (mi + KC - 1) / KC;
    TS reg[EPT];
    auto fetch = [&](int ch) {
      const int64_t base = (int64_t)ch * csz;
#pragma unroll
      for (int u = 0; u < EPT; ++u) {
        const int e = t + NT * u;
        const int64_t g = base + e;
        reg[u] = (e < csz && g < total) ? c.Gw[g] : (TS)0;
      }
    };
    auto stash = [&](int buf) {
      TS* dst = c.Gst + buf * csz;
#pragma unroll
      for (int u = 0; u < EPT; ++u) {
        const int e = t + NT * u;
        if (e < csz) dst[e] = reg[u];
      }
    };
    int offA[SMAX], offB[SMAX];
#pragma unroll
    for (int s = 0; s < SMAX; ++s) {
      offA[s] = 16 * c.sI[s];
      offB[s] = 16 * c.sJ[s];
    }
    __syncthreads();  // vb written; the previous users of the LDS staging area are done
    fetch(0);
    stash(0);
    __syncthreads();
    // fused residual products (gz != nullptr): rows of G x, 16 lanes per row, into vc; the
    // columns of G^T z split over NT / ceil64(n) thread groups, combined after the stream
    const int cpn = (c.n + 63) & ~63, P = NT / cpn > 0 ? NT / cpn : 1, jc = t % cpn, pt = t / cpn;
    double gzacc = 0.0;
    for (int ch = 0; ch < nch; ++ch) {
      if (ch + 1 < nch) fetch(ch + 1);
      const TS* Gc = c.Gst + (ch & 1) * csz;
      if (gz) {
        for (int kb = 4 * c.w; kb < KC; kb += 4 * W) {
          const int kk = kb + rq, k = ch * KC + kk;
          const TS* gr = Gc + kk * ldg + cq;
          double a0 = 0.0;
#pragma unroll
          for (int u = 0; u < NTI; ++u)
            if (u < c.nti) a0 += (double)gr[16 * u] * c.col[16 * u + cq];
          a0 = row16_sum(a0);
          if (cq == 0 && k < mi) c.vc[k] = a0;
        }
        if (pt < P && jc < c.n) {
          for (int kk = pt; kk < KC; kk += P) gzacc += (double)Gc[kk * ldg + jc] * c.vz[ch * KC + kk];
        }
      }
      double a[SMAX], b[SMAX];
      auto ld = [&](int ks, double (&aa)[SMAX], double (&bb)[SMAX]) {
        const int kk = 4 * ks + rq;
        const int k = ch * KC + kk;
        const double dk = k < mi ? c.vb[k] : 0.0;
        const TS* gr = Gc + kk * ldg + cq;
#pragma unroll
        for (int s = 0; s < SMAX; ++s) {
          aa[s] = (double)gr[offA[s]] * dk;
          bb[s] = (double)gr[offB[s]];
        }
      };
      ld(0, a, b);
#pragma unroll
      for (int ks = 0; ks < KC / 4; ++ks) {
        double an[SMAX], bn[SMAX];
        if (ks + 1 < KC / 4) ld(ks + 1, an, bn);
#pragma unroll
        for (int s = 0; s < SMAX; ++s) T[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b[s], T[s], 0, 0, 0);
        if (ks + 1 < KC / 4) {
#pragma unroll
          for (int s = 0; s < SMAX; ++s) { a[s] = an[s]; b[s] = bn[s]; }
        }
      }
      if (ch + 1 < nch) stash((ch + 1) & 1);
      __syncthreads();
    }
    if (gz) {
      if (pt < P) c.part[pt * cpn + jc] = gzacc;
      __syncthreads();
      double v = 0.0;
      if (t < c.n)
        for (int q = 0; q < P; ++q) v += c.part[q * cpn + t];
      *gz = v;
      __syncthreads();
    }
  }
  // reference diagonal: rows >= from of the diagonal tiles
  auto snap_diag = [&](int from) {
#pragma unroll
    for (int s = 0; s < SMAX; ++s) {
      if (c.sv[s] && c.sI[s] == c.sJ[s]) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = 16 * c.sI[s] + cq;
          if (4 * r + rq == cq && row >= from) c.d0[row] = T[s][r];
        }
      }
    }
  };
  snap_diag(0);
  const int Jn = np < m ? (np >> 4) : -1;  // tile column of the first Schur pivot
  for (int Jc = 0; Jc < nti; ++Jc) {
    if (Jc == Jn) snap_diag(np);
    int nq = 0;
    for (int pp = 0; pp < 4; ++pp) {
      const int c0 = 16 * Jc + 4 * pp, jc = 4 * pp;
      if (c0 >= m) break;
      ++nq;
      if (c0 == np && np > 16 * Jc && t >= np && t < m) {
        // the snapshot predates this column's pivots before np: take them out
        double acc = 0.0;
        for (int j = 0; j < np - 16 * Jc; ++j) {
          const double lv = c.pbB[t * LDB + j];
          acc += lv * lv * c.pv[j];
        }
        c.d0[t] -= acc;
      }
      // extract the panel's 4 columns for rows >= 16 Jc
#pragma unroll
      for (int s = 0; s < SMAX; ++s) {
        if (c.sv[s] && c.sJ[s] == Jc) {
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (cq >= jc && cq < jc + 4) c.pbA[(16 * c.sI[s] + 4 * r + rq) * 4 + (cq - jc)] = T[s][r];
        }
      }
      __syncthreads();
      // the diagonal block (rows c0..c0+3) and this thread's row
      double V[4][4], v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int u = 0; u < 4; ++u) V[i][u] = c.pbA[(c0 + i) * 4 + u];
      const int tr = t < rows ? t : rows - 1;
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = c.pbA[tr * 4 + u];
      double dref[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) dref[i] = c.d0[c0 + i];
      double lf[4], pe[4];
#pragma unroll
      for (int tt = 0; tt < 4; ++tt) {
        const int cc = c0 + tt;
        if (cc < m) {
          const double dj = V[tt][tt];
          const bool hp = cc < np;
          const bool ok = hp ? (dj > kPivotRel * fabs(dref[tt])) : (dj < -kPivotRel * fabs(dref[tt]));
          pe[tt] = ok ? dj : (hp ? kHugePivot : -kHugePivot);
          const double inv = recip(pe[tt]);
          if (t == cc) st.rdpiv = inv;
          lf[tt] = (t > cc) ? v[tt] * inv : 0.0;
#pragma unroll
          for (int u = tt + 1; u < 4; ++u) v[u] -= lf[tt] * V[u][tt];
          double lfi[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) lfi[i] = (c0 + i > cc) ? V[i][tt] * inv : 0.0;
#pragma unroll
          for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int u = tt + 1; u < 4; ++u) V[i][u] -= lfi[i] * V[u][tt];
        } else {
          lf[tt] = 0.0;
          pe[tt] = 0.0;
        }
      }
      if (t < m) {
#pragma unroll
        for (int tt = 0; tt < 4; ++tt)
          if (t > c0 + tt) c.Lp[t * (t + 1) / 2 + c0 + tt] = lf[tt];
      }
      if (t < rows) {
#pragma unroll
        for (int tt = 0; tt < 4; ++tt) c.pbB[t * LDB + jc + tt] = lf[tt];
      }
      if (t == 0) {
#pragma unroll
        for (int tt = 0; tt < 4; ++tt) c.pv[jc + tt] = pe[tt];
      }
      __syncthreads();
      // this panel's rank-4 update of the tile column's own tiles
      const double pj = c.pv[jc + rq];
#pragma unroll
      for (int s = 0; s < SMAX; ++s) {
        if (c.sv[s] && c.sJ[s] == Jc) {
          const double aA = -c.pbB[(16 * c.sI[s] + cq) * LDB + jc + rq];
          const double bB = c.pbB[(16 * Jc + cq) * LDB + jc + rq] * pj;
          T[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(aA, bB, T[s], 0, 0, 0);
        }
      }
    }
    // the column's pivots on the tiles right of it
    double pq[4];
#pragma unroll
    for (int kq = 0; kq < 4; ++kq) pq[kq] = c.pv[4 * kq + rq];
#pragma unroll
    for (int s = 0; s < SMAX; ++s) {
      if (c.sv[s] && c.sJ[s] > Jc) {
        const double* ra = c.pbB + (16 * c.sI[s] + cq) * LDB + rq;
        const double* rb = c.pbB + (16 * c.sJ[s] + cq) * LDB + rq;
        double aA[4], bB[4];
#pragma unroll
        for (int kq = 0; kq < 4; ++kq) {
          aA[kq] = kq < nq ? -ra[4 * kq] : 0.0;
          bB[kq] = kq < nq ? rb[4 * kq] * pq[kq] : 0.0;
        }
#pragma unroll
        for (int kq = 0; kq < 4; ++kq) T[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(aA[kq], bB[kq], T[s], 0, 0, 0);
      }
    }
  }
  __syncthreads();
}

// Solve K [dx; dy] = [p1; v4] (qpax/pdip.py:49-63) with v1 = sg*a1 (augmented with v4 = sg*a4 in
// rows np..m-1: the argument pa), v2 = sg*a2, v3 = sg*a3.  Writes st.da and the ds/dz chunks.
template <typename TS, int NTI, int W>
__device__ __forceinline__ void solve(const GCtx<TS, NTI, W>& cin, GSt<mic_of(W)>& st, double pa,
                                      const double (&a2)[mic_of(W)], const double (&a3)[mic_of(W)], double sg,
                                      double (&ods)[mic_of(W)], double (&odz)[mic_of(W)]) {
  const GCtx<TS, NTI, W> c = fresh(cin);
  constexpr int NT = 64 * W, MIC = mic_of(W);
  const int mi = c.mi, m = c.m, t = c.t, w = c.w, l = c.l;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = NT * cc + t;
    if (k < mi) c.vb[k] = st.D[cc] * (sg * a3[cc] - (sg * a2[cc]) / st.z[cc]);
  }
  __syncthreads();
  pa += g_col(c, c.vb);
  if (t >= m) pa = 0.0;
  double* yb = c.col;
  // forward: unit-lower L y = p, one 64-row block at a time
  const int lr = t < m ? t : m - 1;
  const double* Lrow = c.Lp + lr * (lr + 1) / 2;
  for (int bw = 0; bw < W; ++bw) {
    const int j0 = 64 * bw;
    if (j0 >= m) break;
    const int jn = m - j0 < 64 ? m - j0 : 64;
    if (w == bw) {
      // Lanes of rows >= m hold 0 throughout.  The factor entries load 16 at a time ahead of
      // the chain; reads past the end of a row stay inside this problem's LDS image and are
      // discarded by the select.
      double lv[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) lv[u] = Lrow[j0 + u];
#pragma unroll
      for (int jb = 0; jb < 64; jb += 16) {
        double nv[16];
        if (jb + 16 < 64) {
#pragma unroll
          for (int u = 0; u < 16; ++u) nv[u] = Lrow[j0 + jb + 16 + u];
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const int J = j0 + jb + u;
          pa -= (t > J && J < m && t < m ? lv[u] : 0.0) * rl(pa, jb + u);
        }
        if (jb + 16 < 64) {
#pragma unroll
          for (int u = 0; u < 16; ++u) lv[u] = nv[u];
        }
      }
      if (t < m) yb[t] = pa;
    }
    __syncthreads();
    if (w > bw && t < m) {
      double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      int j = 0;
      for (; j + 8 <= jn; j += 8) {
#pragma unroll
        for (int u = 0; u < 8; ++u) a[u] += Lrow[j0 + j + u] * yb[j0 + j + u];
      }
      for (; j < jn; ++j) a[0] += Lrow[j0 + j] * yb[j0 + j];
      pa -= ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    }
  }
  pa = t < m ? pa * st.rdpiv : 0.0;
  // backward: L^T x = y, blocks from the last
  double* xb = c.col + NT;
  for (int bw = W - 1; bw >= 0; --bw) {
    const int j0 = 64 * bw;
    if (j0 >= m) continue;
    const int jn = m - j0 < 64 ? m - j0 : 64;
    if (w == bw) {
      // row J of the factor, entry t (rows clamped to m - 1; reads past the row end stay inside
      // this problem's LDS image and are discarded by the select)
      double lv[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int J = j0 + 63 - u < m ? j0 + 63 - u : m - 1;
        lv[u] = c.Lp[J * (J + 1) / 2 + t];
      }
#pragma unroll
      for (int jb = 0; jb < 64; jb += 16) {
        double nv[16];
        if (jb + 16 < 64) {
#pragma unroll
          for (int u = 0; u < 16; ++u) {
            const int J = j0 + 63 - (jb + 16 + u) < m ? j0 + 63 - (jb + 16 + u) : m - 1;
            nv[u] = c.Lp[J * (J + 1) / 2 + t];
          }
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const int j = 63 - (jb + u);
          const int J = j0 + j;
          if (j > 0) pa -= (t < J && J < m ? lv[u] : 0.0) * rl(pa, j);
        }
        if (jb + 16 < 64) {
#pragma unroll
          for (int u = 0; u < 16; ++u) lv[u] = nv[u];
        }
      }
      if (t < m) xb[t] = pa;
    }
    __syncthreads();
    if (w < bw) {
      double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      int j = 0;
      for (; j + 8 <= jn; j += 8) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int J = j0 + j + u;
          a[u] += c.Lp[J * (J + 1) / 2 + t] * xb[J];
        }
      }
      for (; j < jn; ++j) {
        const int J = j0 + j;
        a[0] += c.Lp[J * (J + 1) / 2 + t] * xb[J];
      }
      pa -= ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    }
  }
  (void)l;
  st.da = t < m ? pa : 0.0;
  __syncthreads();
  if (t < c.ldg) c.col[t] = st.da;
  __syncthreads();
  g_rows(c, c.vc);
  __syncthreads();
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = NT * cc + t;
    if (k < mi) {
      const double v3 = sg * a3[cc], v2 = sg * a2[cc];
      const double dsk = v3 - c.vc[k];
      ods[cc] = dsk;
      odz[cc] = (v2 - st.z[cc] * dsk) / st.s[cc];
    } else {
      ods[cc] = 0.0;
      odz[cc] = 0.0;
    }
  }
}

template <int W, int MIC>
__device__ __forceinline__ double ort(const double (&x)[MIC], const double (&dx)[MIC], int mi, int t, double* red) {
  double a = 1.0;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = 64 * W * cc + t;
    if (k < mi && dx[cc] < 0) a = nmin(a, -x[cc] / dx[cc]);
  }
  return breduce<W>(a, OpMin(), red);
}

template <typename TS, int NTI, int W>
__device__ __forceinline__ void init(const GCtx<TS, NTI, W>& c, GSt<mic_of(W)>& st) {
  constexpr int NT = 64 * W, MIC = mic_of(W);
  const int mi = c.mi, t = c.t;
  factor(c, st, true);
  double zero[MIC];
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) zero[cc] = 0.0;
  solve(c, st, -st.qa, zero, st.h, 1.0, st.ds, st.dz);
  st.xa = st.da;
  double zmax = -INFINITY, zmin = INFINITY;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = NT * cc + t;
    const double zk = -st.ds[cc];
    st.z[cc] = zk;
    if (k < mi) { zmax = nmax(zmax, zk); zmin = nmin(zmin, zk); }
  }
  zmax = breduce<W>(zmax, OpMax(), c.red);
  zmin = breduce<W>(zmin, OpMin(), c.red);
  const double ap = zmax, ad = -zmin;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const double zk = st.z[cc];
    st.s[cc] = (ap < 0) ? -zk : -zk + (1 + ap);
    st.z[cc] = (ad >= 0) ? zk + (1 + ad) : zk;
  }
}

template <typename TS, int NTI, int W>
__device__ __forceinline__ void step(const GCtx<TS, NTI, W>& c, GSt<mic_of(W)>& st, bool relax_rule) {
  constexpr int MIC = mic_of(W);
  const double as = ort<W, MIC>(st.s, st.ds, c.mi, c.t, c.red);
  const double az = ort<W, MIC>(st.z, st.dz, c.mi, c.t, c.red);
  const double alpha = relax_rule ? 0.99 * nmin(nmin(1.0, 0.99 * as), 0.99 * az) : 0.99 * nmin(as, az);
  st.xa += alpha * st.da;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    st.s[cc] += alpha * st.ds[cc];
    st.z[cc] += alpha * st.dz[cc];
  }
}

// One predictor-corrector iteration (qpax/pdip.py:87-119); returns the entry convergence flag.
template <typename TS, int NTI, int W>
__device__ __forceinline__ int pc_iter(const GCtx<TS, NTI, W>& c, GSt<mic_of(W)>& st, double tol) {
  constexpr int NT = 64 * W, MIC = mic_of(W);
  const int mi = c.mi, t = c.t;
  residual_pre(c, st);
  double gz;
  factor(c, st, false, &gz);
  const double nrm = residual_post(c, st, gz, 0.0);
  solve(c, st, -st.ra, st.r2, st.r3, -1.0, st.dsa, st.dza);
  // centering (qpax/pdip.py:76-84)
  double sz = 0.0;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc)
    if (NT * cc + t < mi) sz += st.s[cc] * st.z[cc];
  sz = breduce<W>(sz, OpSum(), c.red);
  const double mu = sz / mi;
  const double a = nmin(ort<W, MIC>(st.s, st.dsa, mi, t, c.red), ort<W, MIC>(st.z, st.dza, mi, t, c.red));
  double gg = 0.0;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc)
    if (NT * cc + t < mi) gg += (st.s[cc] + a * st.dsa[cc]) * (st.z[cc] + a * st.dza[cc]);
  gg = breduce<W>(gg, OpSum(), c.red);
  const double ratio = gg / sz;
  const double sm = ratio * ratio * ratio * mu;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) st.r2[cc] = st.r2[cc] - (sm - st.dsa[cc] * st.dza[cc]);
  solve(c, st, -st.ra, st.r2, st.r3, -1.0, st.ds, st.dz);
  step(c, st, false);
  return nrm < tol ? 1 : 0;
}

// One relaxed Newton iteration (qpax/pdip_relaxed.py:7-52).
template <typename TS, int NTI, int W>
__device__ __forceinline__ int relax_iter(const GCtx<TS, NTI, W>& c, GSt<mic_of(W)>& st, double tol, double kappa) {
  residual_pre(c, st);
  double gz;
  factor(c, st, false, &gz);
  const double nrm = residual_post(c, st, gz, kappa);
  solve(c, st, -st.ra, st.r2, st.r3, -1.0, st.ds, st.dz);
  step(c, st, true);
  return nrm < tol ? 1 : 0;
}

template <typename TS, int NTI, int W>
__device__ __forceinline__ void store_state(const GCtx<TS, NTI, W>& c, const GSt<mic_of(W)>& st, const PIdx& pi, TS* x, TS* s,
                            TS* z, TS* y, bool post) {
  constexpr int NT = 64 * W, MIC = mic_of(W);
  const int n = c.n, np = c.np, mi = c.mi, m = c.m, t = c.t;
  if (t < n) x[pi.oq + t] = (TS)st.xa;
  else if (t >= np && t < m) y[pi.ob + (t - np)] = (TS)st.xa;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = NT * cc + t;
    if (k < mi) {
      double sv = st.s[cc], zv = st.z[cc];
      if (post && st.hinf[cc]) { zv = 0.0; sv = INFINITY; }
      s[pi.oh + k] = (TS)sv;
      z[pi.oh + k] = (TS)zv;
    }
  }
}

template <typename TS>
__device__ __forceinline__ int64_t next_problem_group(const Args<TS>& a, int& round, int* slot) {
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

// Forward kernel: solve (and optionally relax) QPs, one per workgroup at a time.
template <typename TS, int NTI, int W>
__global__ __launch_bounds__(64 * W) void k_group_solve(Args<TS> a, int with_relax) {
  constexpr int MIC = mic_of(W);
  __shared__ int slot;
  int round = 0;
  for (int64_t p; (p = next_problem_group(a, round, &slot)) >= 0;) {
    const PIdx pi = pidx(a, p);
    const auto c = make_gctx<TS, NTI, W>(pi.n, pi.me, pi.mi, a.ws + pi.ows);
    GSt<mic_of(W)> st;
    load(c, st, a.Q + pi.oQ, a.q + pi.oq, a.A + pi.oA, a.b + pi.ob, a.G + pi.oG, a.h + pi.oh);
    init(c, st);
    int conv = 0, it = 0;
    while (it < kMaxIters && conv == 0) {
      conv = pc_iter(c, st, a.tol);
      ++it;
    }
    store_state(c, st, pi, a.x, a.s, a.z, a.y, true);
    if (c.t == 0) { a.conv[p] = conv; a.iters[p] = it; }
    if (!with_relax) continue;
#pragma unroll
    for (int cc = 0; cc < MIC; ++cc)
      if (st.hinf[cc]) { st.z[cc] = 0.0; st.s[cc] = INFINITY; }
    int rconv = 0, rit = 0;
    while (rit < kMaxIters && rconv == 0) {
      rconv = relax_iter(c, st, a.tol, a.kappa);
      ++rit;
    }
    store_state(c, st, pi, a.xr, a.sr, a.zr, a.yr, false);
    if (c.t == 0) { a.rconv[p] = rconv; a.riters[p] = rit; }
  }
}

// Backward kernel (qpax/diff_qp.py:20-46 with optnet_derivatives :8-17).
template <typename TS, int NTI, int W>
__global__ __launch_bounds__(64 * W) void k_group_bwd(Args<TS> a) {
  constexpr int NT = 64 * W, MIC = mic_of(W);
  __shared__ int slot;
  int round = 0;
  for (int64_t p; (p = next_problem_group(a, round, &slot)) >= 0;) {
    const PIdx pi = pidx(a, p);
    const int n = pi.n, me = pi.me, mi = pi.mi;
    const auto c = make_gctx<TS, NTI, W>(n, me, mi, a.ws + pi.ows);
    const int t = c.t, np = c.np, m = c.m, l = c.l, w = c.w;
    GSt<mic_of(W)> st;
    load(c, st, a.Q + pi.oQ, a.q + pi.oq, a.A + pi.oA, a.b + pi.ob, a.G + pi.oG, a.h + pi.oh);
    st.xa = t < n ? (double)a.xr_in[pi.oq + t] : ((t >= np && t < m) ? (double)a.yr_in[pi.ob + (t - np)] : 0.0);
#pragma unroll
    for (int cc = 0; cc < MIC; ++cc) {
      const int k = NT * cc + t;
      st.s[cc] = k < mi ? (double)a.sr_in[pi.oh + k] : 1.0;
      st.z[cc] = k < mi ? (double)a.zr_in[pi.oh + k] : 1.0;
    }
    factor(c, st, false);
    double zero[MIC];
#pragma unroll
    for (int cc = 0; cc < MIC; ++cc) zero[cc] = 0.0;
    const double g = t < n ? (double)a.g[pi.oq + t] : 0.0;
    solve(c, st, -g, zero, zero, 1.0, st.ds, st.dz);
    // cotangents; dz holds dlam_tilde, dlam = dlam_tilde / lam
    __syncthreads();
    double* X = c.col;
    double* DX = c.col + NT;
    if (t < m) { X[t] = st.xa; DX[t] = st.da; }
#pragma unroll
    for (int cc = 0; cc < MIC; ++cc) {
      const int k = NT * cc + t;
      if (k < mi) {
        c.vb[k] = st.z[cc];
        c.vc[k] = st.dz[cc] / st.z[cc];
      }
    }
    __syncthreads();
    for (int i = w; i < n; i += W)
      for (int j = l; j < n; j += 64)
        a.dQ[pi.oQ + (int64_t)i * n + j] = (TS)(0.5 * (DX[i] * X[j] + X[i] * DX[j]));
    if (t < n) a.dq[pi.oq + t] = (TS)st.da;
    for (int ci = w; ci < me; ci += W) {
      const double dnu = DX[np + ci], nu = X[np + ci];
      for (int j = l; j < n; j += 64) a.dA[pi.oA + (int64_t)ci * n + j] = (TS)(dnu * X[j] + nu * DX[j]);
    }
    if (t >= np && t < m) a.db[pi.ob + (t - np)] = (TS)(-st.da);
    const bool exact = a.flags & 1;
    for (int k = w; k < mi; k += W) {
      const double lam = c.vb[k], dlam = c.vc[k];
      for (int j = l; j < n; j += 64) {
        const double v = exact ? lam * dlam * X[j] + lam * DX[j] : lam * (dlam * X[j] + lam * DX[j]);
        a.dG[pi.oG + (int64_t)k * n + j] = (TS)v;
      }
    }
#pragma unroll
    for (int cc = 0; cc < MIC; ++cc) {
      const int k = NT * cc + t;
      if (k < mi) a.dh[pi.oh + k] = (TS)(-st.z[cc] * (st.dz[cc] / st.z[cc]));
    }
    __syncthreads();
  }
}

}  // namespace group
}  // namespace qpx
