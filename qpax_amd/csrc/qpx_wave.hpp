// Fused solver for small dense QPs: one wavefront owns one QP and keeps it on chip.
//
// Rows of the reduced KKT system (lane index l = 0..63):
//   0 .. n-1      x            (the QP's variables)
//   n .. np-1     padding      (np = n rounded up to a multiple of 4; identity rows)
//   np .. m-1     y            (m = np + m_eq <= NMAX <= 64; the equality multipliers)
// The cvxopt reduction of qpax (qpax/pdip.py:38-63) solves
//     K [dx; dy] = [p1; v4],   K = [[H, A^T], [A, 0]],   H = Q + G^T diag(z/s) G.
// K is factored as one quasi-definite LDL^T: pivots 0..np-1 are positive (the Cholesky of H),
// pivots np..m-1 negative (minus the Schur complement F = A H^{-1} A^T) -- the reference's
// H-then-F factorisation done as a single elimination.
//
// Registers hold K in the float64 matrix-core tile layout (16x16 tiles of the lower triangle,
// element (16I + 4r + l/16, 16J + l%16) of tile (I, J) in register r of lane l), the augmented
// vectors [x; y] one element per lane, and the m_ineq-vectors in MIC chunks of 64 rows (row
// 64c + l in lane l).  G^T D G accumulates into the tiles with v_mfma_f64_16x16x4; the
// elimination runs in panels of 4 pivots whose rank-4 trailing update is one such MFMA per tile.
// LDS holds G (input precision), the packed unit-lower factor (row l at l(l+1)/2) for the two
// triangular sweeps, and a panel/broadcast buffer.
#pragma once
#include "qpx_args.hpp"
#include "qpx_core.hpp"
#include "qpx_wave_launch.hpp"

namespace qpx {
namespace wave {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double rl(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// Wavefront reductions: butterfly inside each row of 16 lanes with DPP (quad_perm [1,0,3,2],
// quad_perm [2,3,0,1], row_half_mirror, row_mirror), then the four row results through readlane.
// Every lane receives the result.
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
template <class Op>
__device__ __forceinline__ double wreduce(double v, Op op) {
  v = op(v, dpp<0xB1>(v));   // quad_perm [1,0,3,2]
  v = op(v, dpp<0x4E>(v));   // quad_perm [2,3,0,1]
  v = op(v, dpp<0x141>(v));  // row_half_mirror
  v = op(v, dpp<0x140>(v));  // row_mirror
  return op(op(rl(v, 0), rl(v, 16)), op(rl(v, 32), rl(v, 48)));
}
__device__ __forceinline__ double wmax(double v) { return wreduce(v, OpMax()); }
__device__ __forceinline__ double wmin(double v) { return wreduce(v, OpMin()); }
__device__ __forceinline__ double wsum(double v) { return wreduce(v, OpSum()); }
// Sum over the 16 lanes of each DPP row; every lane of the row receives it.
__device__ __forceinline__ double row16_sum(double v) {
  v += dpp<0xB1>(v);
  v += dpp<0x4E>(v);
  v += dpp<0x141>(v);
  v += dpp<0x140>(v);
  return v;
}

#define WSYNC() __builtin_amdgcn_wave_barrier()

// sum_k G[k][col] * v[k] over k < mi (v broadcast from LDS); eight rows per step, with two
// buffers: the next step's loads are in flight while this step accumulates (a last step reloads
// its own rows).
template <typename TS>
__device__ __forceinline__ double gcol_dot(const TS* __restrict__ Gs, int n, int mi, int col,
                                           const double* __restrict__ v) {
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  const int nk = mi & ~7;
  auto ld = [&](int k, TS (&g)[8], double2 (&w)[4]) {
#pragma unroll
    for (int u = 0; u < 8; ++u) g[u] = Gs[(k + u) * n + col];
#pragma unroll
    for (int u = 0; u < 4; ++u) w[u] = *reinterpret_cast<const double2*>(v + k + 2 * u);
  };
  auto acc = [&](const TS (&g)[8], const double2 (&w)[4]) {
#pragma unroll
    for (int u = 0; u < 8; u += 2) {
      a[(u >> 1) & 3] += (double)g[u] * w[u >> 1].x;
      a[(u >> 1) & 3] += (double)g[u + 1] * w[u >> 1].y;
    }
  };
  if (nk > 0) {
    TS g0[8], g1[8];
    double2 w0[4], w1[4];
    ld(0, g0, w0);
    for (int k = 0;;) {
      ld(k + 8 < nk ? k + 8 : k, g1, w1);
      __builtin_amdgcn_sched_barrier(0);
      acc(g0, w0);
      k += 8;
      if (k >= nk) break;
      ld(k + 8 < nk ? k + 8 : k, g0, w0);
      __builtin_amdgcn_sched_barrier(0);
      acc(g1, w1);
      k += 8;
      if (k >= nk) break;
    }
  }
  for (int k = nk; k < mi; ++k) a[0] += (double)Gs[k * n + col] * v[k];
  return (a[0] + a[1]) + (a[2] + a[3]);
}

// sum_j G[row][j] * v[j] over j < n (v broadcast from LDS); four columns per trip.
template <typename TS>
__device__ __forceinline__ double grow_dot(const TS* __restrict__ Gs, int n, int row, const double* __restrict__ v) {
  const TS* g = Gs + row * n;
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  int j = 0;
  for (; j + 8 <= n; j += 8) {
    double gg[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) gg[u] = (double)g[j + u];
#pragma unroll
    for (int u = 0; u < 8; u += 2) {
      const double2 vv = *reinterpret_cast<const double2*>(v + j + u);
      a[(u >> 1) & 3] += gg[u] * vv.x;
      a[(u >> 1) & 3] += gg[u + 1] * vv.y;
    }
  }
  for (; j < n; ++j) a[0] += (double)g[j] * v[j];
  return (a[0] + a[1]) + (a[2] + a[3]);
}

// grow_dot for even n: the row starts on a two-element boundary, so entries load in pairs; two
// buffers as in gcol_dot.
template <typename TS>
__device__ __forceinline__ double grow_dot2(const TS* __restrict__ Gs, int n, int row, const double* __restrict__ v) {
  typedef TS V2 __attribute__((ext_vector_type(2)));
  const V2* g = reinterpret_cast<const V2*>(Gs + row * n);
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  const int n2 = n >> 1;
  const int nj = n2 & ~3;
  auto ld = [&](int j, V2 (&gg)[4], double2 (&w)[4]) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      gg[u] = g[j + u];
      w[u] = *reinterpret_cast<const double2*>(v + 2 * (j + u));
    }
  };
  auto acc = [&](const V2 (&gg)[4], const double2 (&w)[4]) {
#pragma unroll
    for (int u = 0; u < 4; ++u) a[u] += (double)gg[u].x * w[u].x + (double)gg[u].y * w[u].y;
  };
  if (nj > 0) {
    V2 g0[4], g1[4];
    double2 w0[4], w1[4];
    ld(0, g0, w0);
    for (int j = 0;;) {
      ld(j + 4 < nj ? j + 4 : j, g1, w1);
      __builtin_amdgcn_sched_barrier(0);
      acc(g0, w0);
      j += 4;
      if (j >= nj) break;
      ld(j + 4 < nj ? j + 4 : j, g0, w0);
      __builtin_amdgcn_sched_barrier(0);
      acc(g1, w1);
      j += 4;
      if (j >= nj) break;
    }
  }
  for (int j = nj; j < n2; ++j) {
    const V2 gg = g[j];
    a[0] += (double)gg.x * v[2 * j] + (double)gg.y * v[2 * j + 1];
  }
  return (a[0] + a[1]) + (a[2] + a[3]);
}

// Box rows -- the +-I blocks of box-constrained and nonnegative QPs (G = -I in the NNLS batch
// example, G = [G_dense; I; -I] in the reference's QPTEST).  load() finds md, one past the last
// row of G that is not +-e_j or zero; rows md..mi-1 are each +-e_j (or zero, e.g. a row with
// infinite h), and their code sigma * (j + 1) (0 for a zero row) replaces the row's first entry in
// LDS.  The dense streams stop at md, and those rows enter as a diagonal (G^T D G), a gather
// (G^T v) and one entry (G x).
template <typename TS>
__device__ __forceinline__ void box_code(const TS* __restrict__ Gs, int n, int k, int& j, double& sg) {
  const double cv = (double)Gs[k * n];
  sg = cv > 0.0 ? 1.0 : (cv < 0.0 ? -1.0 : 0.0);
  j = (int)fabs(cv) - 1;
}
// sum of sigma_k^P v[k] over the box rows k in [md, mi) whose column j_k is col (P = 1: G^T v;
// P = 2: the diagonal of G^T diag(v) G)
template <int P, typename TS>
__device__ __forceinline__ double box_gather(const TS* __restrict__ Gs, int n, int md, int mi, int col,
                                             const double* __restrict__ v) {
  double a = 0.0;
  for (int k = md; k < mi; ++k) {
    int j;
    double sg;
    box_code(Gs, n, k, j, sg);
    a += j == col ? (P == 2 ? sg * sg : sg) * v[k] : 0.0;
  }
  return a;
}
// (G x)_k of box row k, x broadcast from LDS
template <typename TS>
__device__ __forceinline__ double box_row(const TS* __restrict__ Gs, int n, int k, const double* __restrict__ x) {
  int j;
  double sg;
  box_code(Gs, n, k, j, sg);
  return sg * x[j < 0 ? 0 : j];
}

// 1/d to full double precision: hardware reciprocal plus two Newton steps.
__device__ __forceinline__ double recip(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  r = fma(r, fma(-d, r, 1.0), r);
  return r;
}

template <bool B>
struct BoolTag {
  static constexpr bool value = B;
};

// 1/d with a shorter dependency chain: the hardware reciprocal r0 is accurate to 2^29 ulp
// (|e| <= 2^-23 with e = 1 - d r0), and 1/d = r0 (1 + e + e^2 + ...), so r0 + r0 (e + e^2) leaves
// a truncation of e^3 <= 2^-69 under the rounding of the three fused operations.
__device__ __forceinline__ double recip_c(double d) {
  const double r = __builtin_amdgcn_rcp(d);
  const double e = fma(-d, r, 1.0);
  return fma(r, fma(e, e, e), r);
}


template <typename TS, int NMAX, int MIC>
struct St {
  int md;         // rows of G streamed densely; rows md..mi-1 are box rows (see box_code)
  double qa;      // q (rows < n), -b (rows np..m-1), 0 otherwise
  double h[MIC];
  bool hinf[MIC];
  double rdpiv;   // 1 / pivot of this lane's row
  double xa, ra, da;
  double s[MIC], z[MIC], D[MIC], r2[MIC], r3[MIC], dsa[MIC], dza[MIC], ds[MIC], dz[MIC];
  // the next factorisation's initial tiles (K's constant part, upper triangle), loaded a
  // factorisation ahead so that the loads complete behind the solves
  d4 Tk[NMAX / 16 * (NMAX / 16 + 1) / 2];
};

template <typename TS, int NMAX, int MIC>
struct Ctx {
  int n, np, me, mi, m, l;
  // K's constant part [[Q_sym, A^T], [A, 0]] (m x m, float64, symmetric, identity on the padding
  // rows) in this problem's global workspace; entry (i, j) at KQ[j*m + i].
  double* KQ;
  const TS* Gs;   // G, row-major mi x n
  double* Lp;     // packed unit-lower factor
  double* pb;     // panel buffer (64 x 8 doubles); outside the factorisation: col + vb
  double* col;    // 128 doubles (aliases pb)
  double* vb;     // m_ineq doubles (aliases pb + 128)
};

// K's constant part as the upper-triangle tiles of factor_up (tile (J, I), J <= I, element
// (16J + 4r + l/16, 16I + l%16) in register r), read along its rows (coalesced over the lanes).
template <typename TS, int NMAX, int MIC>
__device__ __forceinline__ void tile_load(const Ctx<TS, NMAX, MIC>& c, d4 (&T)[NMAX / 16 * (NMAX / 16 + 1) / 2]) {
  constexpr int NTI = NMAX / 16;
  const int m = c.m, l = c.l, rq = l >> 4, cq = l & 15;
#pragma unroll
  for (int I = 0; I < NTI; ++I) {
#pragma unroll
    for (int J = 0; J <= I; ++J) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = 16 * J + 4 * r + rq, cl = 16 * I + cq;
        const bool in = row < m && cl < m;
        const double v = c.KQ[(in ? row : 0) * m + (in ? cl : 0)];
        T[I * (I + 1) / 2 + J][r] = in ? v : 0.0;
      }
    }
  }
}

template <bool FAST, typename TS, int NMAX, int MIC>
__device__ __forceinline__ void load(const Ctx<TS, NMAX, MIC>& c, St<TS, NMAX, MIC>& st, const TS* Q,
                                     const TS* q, const TS* A, const TS* b, const TS* G, const TS* h,
                                     bool build_kq = true) {
  const int n = c.n, np = c.np, mi = c.mi, m = c.m, l = c.l, me = c.me;
  TS* Gs = const_cast<TS*>(c.Gs);
  // G into LDS as one flat copy.  FAST with 16-byte-aligned source and destination and at most
  // 32 vectors per lane: every load is issued first (indices clamped, so no lane-divergent
  // branch), the constant-part build below runs while they are in flight, and the stores follow
  // it.  Otherwise: 16-byte vectors 16 in flight per lane (FAST), or 8 scalar loads in flight.
  typedef TS V __attribute__((ext_vector_type(16 / sizeof(TS))));
  constexpr int VW = 16 / sizeof(TS);
  constexpr int GV = 32;
  const bool vec = FAST && ((reinterpret_cast<uintptr_t>(G) | reinterpret_cast<uintptr_t>(Gs)) & 15) == 0;
  const int gtot = mi * n, gnv = gtot / VW, gnu = (gnv + 63) / 64;
  const bool gearly = vec && gnu <= GV;
  V gbuf[GV];
  if (gearly) {
    const V* Gv = reinterpret_cast<const V*>(G);
#pragma unroll
    for (int u = 0; u < GV; ++u) {
      const int e = 64 * u + l;
      if (u < gnu) gbuf[u] = Gv[e < gnv ? e : gnv - 1];
    }
  } else if (vec) {
    const V* Gv = reinterpret_cast<const V*>(G);
    V* Gsv = reinterpret_cast<V*>(Gs);
    for (int e0 = 0; e0 < gnv; e0 += 16 * 64) {
      V g[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int e = e0 + 64 * u + l;
        if (e < gnv) g[u] = Gv[e];
      }
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int e = e0 + 64 * u + l;
        if (e < gnv) Gsv[e] = g[u];
      }
    }
    for (int e = gnv * VW + l; e < gtot; e += 64) Gs[e] = G[e];
  } else {
    int e = l;
    for (; e + 7 * 64 < gtot; e += 8 * 64) {
      TS g[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) g[u] = G[e + 64 * u];
#pragma unroll
      for (int u = 0; u < 8; ++u) Gs[e + 64 * u] = g[u];
    }
    for (; e < gtot; e += 64) Gs[e] = G[e];
  }
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = 64 * cc + l;
    const double hk = k < mi ? (double)h[k] : 0.0;
    st.hinf[cc] = isinf(hk);
    st.h[cc] = st.hinf[cc] ? 0.0 : hk;
    st.s[cc] = 1.0;
    st.z[cc] = 1.0;
  }
  // rows with infinite h: diag(1 * ~mask) @ G (pdip.py:176), 0 * g (after G is in LDS)
  auto zero_inf_rows = [&]() {
    WSYNC();
#pragma unroll
    for (int cc = 0; cc < MIC; ++cc) {
      uint64_t mask = __ballot(st.hinf[cc]);
      while (mask) {
        const int k = 64 * cc + __builtin_ctzll(mask);
        mask &= mask - 1;
        for (int j = l; j < n; j += 64) Gs[k * n + j] = (TS)0 * Gs[k * n + j];
      }
    }
  };
  if (!gearly) zero_inf_rows();
  // K's constant part.  FAST: every lane issues the four candidate loads of each element at
  // clamped (always valid) addresses -- Q[l][j], Q[j][l], A[j - np][l], A[l - np][j] -- for 16
  // columns at a time, then selects, so no load waits behind a lane-divergent branch.
  if constexpr (FAST) {
    if (l < m && build_kq) {
      constexpr int CB = 16;
      const int lq = l < n ? l : n - 1;
      const int la = me > 0 ? (l < np ? 0 : (l - np < me ? l - np : me - 1)) : 0;
      for (int j0 = 0; j0 < m; j0 += CB) {
        TS q1[CB], q2[CB], a1[CB], a2[CB];
#pragma unroll
        for (int u = 0; u < CB; ++u) {
          const int j = j0 + u, jq = j < n ? j : n - 1;
          q1[u] = Q[lq * n + jq];
          q2[u] = Q[jq * n + lq];
        }
        if (me > 0) {
#pragma unroll
          for (int u = 0; u < CB; ++u) {
            const int j = j0 + u, jq = j < n ? j : n - 1;
            const int ja = j < np ? 0 : (j - np < me ? j - np : me - 1);
            a1[u] = A[ja * n + lq];
            a2[u] = A[la * n + jq];
          }
        }
#pragma unroll
        for (int u = 0; u < CB; ++u) {
          const int j = j0 + u;
          if (j >= m) break;
          double val;
          if (l < n) {
            val = j < n ? 0.5 * ((double)q1[u] + (double)q2[u]) : (j >= np && me > 0 ? (double)a1[u] : 0.0);
          } else if (l < np) {
            val = j == l ? 1.0 : 0.0;
          } else {
            val = j < n ? (double)a2[u] : 0.0;
          }
          c.KQ[j * m + l] = val;
        }
      }
    }
  } else {
  constexpr int CB = 8;
  if (l < m) {
    for (int j0 = 0; j0 < m; j0 += CB) {
      double v[CB];
#pragma unroll
      for (int u = 0; u < CB; ++u) {
        const int j = j0 + u;
        double val = 0.0;
        if (j < m) {
          if (l < n) {
            if (j < n) val = 0.5 * ((double)Q[l * n + j] + (double)Q[j * n + l]);
            else if (j >= np) val = (double)A[(j - np) * n + l];
          } else if (l < np) {
            val = (j == l) ? 1.0 : 0.0;
          } else if (j < n) {
            val = (double)A[(l - np) * n + j];
          }
        }
        v[u] = val;
      }
#pragma unroll
      for (int u = 0; u < CB; ++u)
        if (j0 + u < m) c.KQ[(j0 + u) * m + l] = v[u];
    }
  }
  }
  if (gearly) {
    V* Gsv = reinterpret_cast<V*>(Gs);
#pragma unroll
    for (int u = 0; u < GV; ++u) {
      const int e = 64 * u + l;
      if (u < gnu) Gsv[e < gnv ? e : gnv - 1] = gbuf[u];
    }
    for (int e = gnv * VW + l; e < gtot; e += 64) Gs[e] = G[e];
    zero_inf_rows();
  }
  // box rows: md = one past the last row that is not +-e_j or zero; rows md..mi-1 get their codes
  WSYNC();
  {
    double last = 0.0, code[MIC];
#pragma unroll
    for (int cc = 0; cc < MIC; ++cc) {
      const int k = 64 * cc + l;
      code[cc] = 0.0;
      if (k < mi) {
        // a row with a second nonzero is dense: general rows end the scan after two entries
        int nnz = 0, jj = 0;
        double v1 = 0.0;
        for (int j = 0; j < n && nnz < 2; ++j) {
          const double g = (double)Gs[k * n + j];
          if (g != 0.0) {
            ++nnz;
            jj = j;
            v1 = g;
          }
        }
        const bool box = nnz == 0 || (nnz == 1 && (v1 == 1.0 || v1 == -1.0));
        code[cc] = nnz == 0 ? 0.0 : (v1 > 0.0 ? 1.0 : -1.0) * (jj + 1);
        if (!box) last = (double)(k + 1);
      }
    }
    st.md = (int)wmax(last);
    WSYNC();
#pragma unroll
    for (int cc = 0; cc < MIC; ++cc) {
      const int k = 64 * cc + l;
      if (k >= st.md && k < mi) Gs[k * n] = (TS)code[cc];
    }
  }
  __threadfence_block();
  if constexpr (FAST) tile_load(c, st.Tk);
  st.qa = l < n ? (double)q[l] : ((l >= np && l < m) ? -(double)b[l - np] : 0.0);
  st.xa = 0.0;
  WSYNC();
}

// Residuals of the (relaxed) KKT system (qpax/pdip.py:93-100); returns the infinity norm
// (NaN-propagating).
template <typename TS, int NMAX, int MIC>
__device__ __forceinline__ double residuals(const Ctx<TS, NMAX, MIC>& c, St<TS, NMAX, MIC>& st, double kappa) {
  const int n = opaque_s(c.n), mi = opaque_s(c.mi), m = opaque_s(c.m), l = opaque_v(c.l), md = opaque_s(st.md);
  c.col[l] = st.xa;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = 64 * cc + l;
    if (k < mi) c.vb[k] = st.z[cc];
  }
  WSYNC();
  // K's constant part times [x; y]: column j of KQ is read coalesced across the lanes, 16 loads in
  // flight per trip (rows >= m read row m - 1 and are discarded)
  double acc = st.qa;
  {
    const double* kcol = c.KQ + (l < m ? l : m - 1);
    double a[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) a[u] = 0.0;
    int j = 0;
    for (; j + 16 <= m; j += 16) {
      double kv[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) kv[u] = kcol[(j + u) * m];
#pragma unroll
      for (int u = 0; u < 16; ++u) a[u] += kv[u] * c.col[j + u];
    }
    for (; j < m; ++j) a[0] += kcol[j * m] * c.col[j];
#pragma unroll
    for (int u = 1; u < 16; ++u) a[0] += a[u];
    if (l < m) acc += a[0];
  }
  if (l < n) acc += gcol_dot(c.Gs, n, md, l, c.vb) + (md < mi ? box_gather<1>(c.Gs, n, md, mi, l, c.vb) : 0.0);
  st.ra = l < m ? acc : 0.0;
  double nrm = fabs(st.ra);
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = 64 * cc + l;
    if (k < mi) {
      const double g = k < md ? grow_dot(c.Gs, n, k, c.col) : box_row(c.Gs, n, k, c.col);
      st.r3[cc] = g + st.s[cc] - st.h[cc];
      st.r2[cc] = st.s[cc] * st.z[cc] - kappa;
      nrm = nmax(nrm, nmax(fabs(st.r2[cc]), fabs(st.r3[cc])));
    } else {
      st.r3[cc] = 0.0;
      st.r2[cc] = 0.0;
    }
  }
  WSYNC();
  return wmax(nrm);
}

__host__ __device__ constexpr int tidx(int I, int J) { return I * (I + 1) / 2 + J; }

// The box rows' part of a factorisation pass (no-op without box rows): D_k on the diagonal of the
// tiles (diagonal tile J holds (16J + c, 16J + c) in register c / 4 of lane 16 (c % 4) + c, the
// same in the lower- and the upper-triangle layouts), and with RES their entries of G x (into vb,
// after the D values there are consumed) and this lane's column of their G^T z (returned).
template <typename TS, int NMAX, int MIC, int NTILE>
__device__ __forceinline__ double box_rows_into(const Ctx<TS, NMAX, MIC>& c, const St<TS, NMAX, MIC>& st,
                                                d4 (&T)[NTILE], int md, bool res) {
  constexpr int NTI = NMAX / 16;
  const int n = c.n, mi = c.mi, l = c.l, rq = l >> 4, cq = l & 15;
  if (md >= mi) return 0.0;
  const double dsum = box_gather<2>(c.Gs, n, md, mi, l, c.vb);
  const bool holds = rq == (cq & 3);
#pragma unroll
  for (int J = 0; J < NTI; ++J) {
    const double v = __shfl(dsum, 16 * J + cq, 64);
#pragma unroll
    for (int r = 0; r < 4; ++r) T[tidx(J, J)][r] += (holds && (cq >> 2) == r) ? v : 0.0;
  }
  double zb = 0.0;
  if (res) {
    zb = l < n ? box_gather<1>(c.Gs, n, md, mi, l, c.vb + mi) : 0.0;
    WSYNC();
#pragma unroll
    for (int cc = 0; cc < MIC; ++cc) {
      const int k = 64 * cc + l;
      if (k >= md && k < mi) c.vb[k] = box_row(c.Gs, n, k, c.col);
    }
    WSYNC();
  }
  return zb;
}

// Factorisation (qpax/pdip.py:38-46): D = z/s (or 1 for the initialisation), K = KQ + G^T D G,
// K = L D L^T with L unit lower (rows packed into c.Lp) and 1/D in st.rdpiv.  A pivot that is not
// beyond kPivotRel times its reference diagonal -- the diagonal entry before elimination for the
// H block, the diagonal entry once all H pivots are done for the Schur block -- or is NaN is
// replaced by kHugePivot with the block's sign, which zeroes that component of later solves.
//
// With RES the same pass forms the residuals of the (relaxed) KKT system at the current point
// (qpax/pdip.py:93-100): K's constant part times [x; y] from the tiles before G^T D G enters
// them, and G x and G^T z from the G operands the G^T D G stream loads anyway (vector ALU work
// under the matrix-core work).  Returns their infinity norm (NaN-propagating), 0 without RES.
template <bool RES, typename TS, int NMAX, int MIC>
__device__ __forceinline__ double factor(const Ctx<TS, NMAX, MIC>& c, St<TS, NMAX, MIC>& st, bool unit_d,
                                         double kappa) {
  constexpr int NTI = NMAX / 16;
  constexpr int NTILE = NTI * (NTI + 1) / 2;
  const int n = opaque_s(c.n), np = opaque_s(c.np), mi = opaque_s(c.mi), m = opaque_s(c.m), l = opaque_v(c.l);
  const int md = opaque_s(st.md);
  const int rq = l >> 4, cq = l & 15;
  double* vz = c.vb + mi;  // z, for G^T z
  double* kx = c.col + 64; // K's constant part times [x; y], then plus G^T z
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = 64 * cc + l;
    st.D[cc] = unit_d ? 1.0 : st.z[cc] / st.s[cc];
    if (k < mi) {
      c.vb[k] = st.D[cc];
      if (RES) vz[k] = st.z[cc];
    }
  }
  if (RES) c.col[l] = st.xa;
  // tiles start as K's constant part (symmetric; with RES read along its rows, coalesced over the
  // lanes)
  d4 T[NTILE];
#pragma unroll
  for (int I = 0; I < NTI; ++I) {
#pragma unroll
    for (int J = 0; J <= I; ++J) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = 16 * I + 4 * r + rq, cl = 16 * J + cq;
        const bool in = row < m && cl < m;
        const double v = RES ? c.KQ[(in ? row : 0) * m + (in ? cl : 0)] : c.KQ[(in ? cl : 0) * m + (in ? row : 0)];
        T[tidx(I, J)][r] = in ? v : 0.0;
      }
    }
  }
  WSYNC();
  // [x; y] at this lane's tile columns (xc) and tile rows (xr)
  double xc[NTI], pc[NTI], zacc[NTI];
#pragma unroll
  for (int J = 0; J < NTI; ++J) {
    xc[J] = RES ? c.col[16 * J + cq] : 0.0;
    pc[J] = 0.0;
    zacc[J] = 0.0;
  }
  if (RES) {
    // K's constant part times [x; y]: the stored lower tiles act directly on their rows and, off
    // the diagonal, transposed on their columns.  Row sums over the 16 lanes of a DPP row, column
    // sums over the four lane groups of 16.
    double pr[NTI][4];
#pragma unroll
    for (int I = 0; I < NTI; ++I) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double xr = c.col[16 * I + 4 * r + rq];
        double a = 0.0;
#pragma unroll
        for (int J = 0; J <= I; ++J) {
          a += T[tidx(I, J)][r] * xc[J];
          if (J < I) pc[J] += T[tidx(I, J)][r] * xr;
        }
        pr[I][r] = a;
      }
    }
#pragma unroll
    for (int I = 0; I < NTI; ++I) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double v = row16_sum(pr[I][r]);
        if (cq == 0) kx[16 * I + 4 * r + rq] = v;
      }
    }
  }
  // G^T D G into the tiles, four rows of G per step, the next step's operands loaded while the
  // current step's MFMAs run.  Columns >= n load as exact zeros.  No branch may separate the
  // MFMAs (the accumulators would be copied around each one).  With RES: the rows of G x (sum
  // over the 16 lanes of a DPP row, stored over the D entries already consumed) and this lane's
  // share of G^T z.
  {
    int off[NTI];
    double msk[NTI];
#pragma unroll
    for (int I = 0; I < NTI; ++I) {
      const int cl = 16 * I + cq;
      off[I] = cl < n ? cl : n - 1;
      msk[I] = cl < n ? 1.0 : 0.0;
    }
    auto fetch = [&](int k0, double (&g)[NTI], double& dk, double& zk) {
      const int k = k0 + rq;
      const int kc = k < md ? k : (md > 0 ? md - 1 : 0);
      dk = k < md ? c.vb[kc] : 0.0;
      zk = (RES && k < md) ? vz[kc] : 0.0;
      const TS* grow = c.Gs + kc * n;
#pragma unroll
      for (int I = 0; I < NTI; ++I) g[I] = (double)grow[off[I]] * msk[I];
    };
    double gc[NTI], dc, zc;
    fetch(0, gc, dc, zc);
    for (int k0 = 0; k0 < md; k0 += 4) {
      double gn[NTI], dn, zn;
      fetch(k0 + 4 < md ? k0 + 4 : k0, gn, dn, zn);
#pragma unroll
      for (int I = 0; I < NTI; ++I) {
        const double a = gc[I] * dc;
#pragma unroll
        for (int J = 0; J <= I; ++J)
          T[tidx(I, J)] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, gc[J], T[tidx(I, J)], 0, 0, 0);
      }
      if (RES) {
        double gx = 0.0;
#pragma unroll
        for (int I = 0; I < NTI; ++I) {
          gx += gc[I] * xc[I];
          zacc[I] += gc[I] * zc;
        }
        gx = row16_sum(gx);
        const int k = k0 + rq;
        if (cq == 0 && k < md) c.vb[k] = gx;
      }
#pragma unroll
      for (int I = 0; I < NTI; ++I) gc[I] = gn[I];
      dc = dn;
      zc = zn;
    }
  }
  WSYNC();
  const double zbox = box_rows_into(c, st, T, md, RES);
  double nrm = 0.0;
  if (RES) {
    // columns of the off-diagonal tiles and of G^T z: sum over the lane groups, add to kx
#pragma unroll
    for (int J = 0; J < NTI; ++J) {
      double v = pc[J] + zacc[J];
      v += __shfl_xor(v, 16, 64);
      v += __shfl_xor(v, 32, 64);
      pc[J] = v;
    }
    if (rq == 0) {
#pragma unroll
      for (int J = 0; J < NTI; ++J) kx[16 * J + cq] += pc[J];
    }
    WSYNC();
    st.ra = l < m ? kx[l] + st.qa + zbox : 0.0;
    nrm = fabs(st.ra);
#pragma unroll
    for (int cc = 0; cc < MIC; ++cc) {
      const int k = 64 * cc + l;
      if (k < mi) {
        st.r3[cc] = c.vb[k] + st.s[cc] - st.h[cc];
        st.r2[cc] = st.s[cc] * st.z[cc] - kappa;
        nrm = nmax(nrm, nmax(fabs(st.r2[cc]), fabs(st.r3[cc])));
      } else {
        st.r3[cc] = 0.0;
        st.r2[cc] = 0.0;
      }
    }
    WSYNC();
    nrm = wmax(nrm);
  }
  double* pb = c.pb;
  // Reference diagonal of this lane's row: the entry before elimination.  Schur rows (l >= np)
  // measure their pivots against the entry once all H pivots are done, i.e. minus the sum of
  // lf^2 * pivot over the H pivots, accumulated in hsum as the pivots go by.
#pragma unroll
  for (int I = 0; I < NTI; ++I) {
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (4 * r + rq == cq) pb[16 * I + cq] = T[tidx(I, I)][r];
  }
  WSYNC();
  const double dref = pb[l];
  WSYNC();
  double hsum = 0.0;
  // Tile columns of 16 pivots, each as 4 panels of 4.  A panel's rank-4 update reaches the
  // column's own tiles at once; the tiles right of the column receive its 16 pivots as one
  // rank-16 update (four MFMAs per tile) when the column is done, with the multipliers read back
  // from the packed factor and the pivots from its diagonal slots (which the sweeps never read).
  // Pivots c >= m meet zero rows and columns: they fall below the floor and only ever scale
  // zeros.
#pragma unroll
  for (int J = 0; J < NTI; ++J) {
    if (16 * J >= m) break;
#pragma unroll
    for (int pp = 0; pp < 4; ++pp) {
      const int c0 = 16 * J + 4 * pp, jc = 4 * pp;
      // floor of each pivot of the panel: kPivotRel times its row's reference diagonal
      const double rf = kPivotRel * fabs(l >= np ? dref - hsum : dref);
      double thr[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) thr[t] = rl(rf, c0 + t);
      // extract the panel's 4 columns, one row per lane
#pragma unroll
      for (int I = 0; I < NTI; ++I) {
        if (I < J) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (cq >= jc && cq < jc + 4) pb[(16 * I + 4 * r + rq) * 4 + (cq - jc)] = T[tidx(I, J)][r];
      }
      WSYNC();
      // this lane's row and the 4 x 4 diagonal block, which every lane eliminates for itself
      double v[4], V[4][4];
      {
        const double2 v01 = *reinterpret_cast<const double2*>(pb + l * 4);
        const double2 v23 = *reinterpret_cast<const double2*>(pb + l * 4 + 2);
        v[0] = v01.x; v[1] = v01.y; v[2] = v23.x; v[3] = v23.y;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const double2 a01 = *reinterpret_cast<const double2*>(pb + (c0 + i) * 4);
          const double2 a23 = *reinterpret_cast<const double2*>(pb + (c0 + i) * 4 + 2);
          V[i][0] = a01.x; V[i][1] = a01.y; V[i][2] = a23.x; V[i][3] = a23.y;
        }
      }
      WSYNC();
      double lf[4], pe[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int cc = c0 + t;
        const double sgn = cc < np ? 1.0 : -1.0;
        const double dj = V[t][t];
        const bool ok = sgn * dj > thr[t];
        pe[t] = ok ? dj : sgn * kHugePivot;
        const double inv = recip(pe[t]);
        if (l == cc) st.rdpiv = inv;
        lf[t] = (l > cc) ? v[t] * inv : 0.0;
#pragma unroll
        for (int u = t + 1; u < 4; ++u) v[u] -= lf[t] * V[u][t];
#pragma unroll
        for (int i = t + 1; i < 4; ++i) {
          const double li = V[i][t] * inv;
#pragma unroll
          for (int u = t + 1; u <= i; ++u) V[i][u] -= li * V[u][t];
        }
        if (cc < np) hsum += lf[t] * lf[t] * pe[t];
      }
      // multipliers into the packed factor, pivots into its diagonal slots
#pragma unroll
      for (int t = 0; t < 4; ++t)
        if (l > c0 + t && l < m) c.Lp[l * (l + 1) / 2 + c0 + t] = lf[t];
      if (l < 4 && c0 + l < m) {
        const double pl = l == 0 ? pe[0] : l == 1 ? pe[1] : l == 2 ? pe[2] : pe[3];
        c.Lp[(c0 + l) * (c0 + l + 3) / 2] = pl;
      }
      *reinterpret_cast<double2*>(pb + l * 8) = double2{lf[0], lf[1]};
      *reinterpret_cast<double2*>(pb + l * 8 + 2) = double2{lf[2], lf[3]};
      *reinterpret_cast<double2*>(pb + l * 8 + 4) = double2{lf[0] * pe[0], lf[1] * pe[1]};
      *reinterpret_cast<double2*>(pb + l * 8 + 6) = double2{lf[2] * pe[2], lf[3] * pe[3]};
      WSYNC();
      // rank-4 update of the column's own tiles (the last panel of a column has none left)
      if (pp < 3) {
        const double bJ = pb[(16 * J + cq) * 8 + 4 + rq];
        double aA[NTI];
#pragma unroll
        for (int I = 0; I < NTI; ++I) aA[I] = (I >= J) ? -pb[(16 * I + cq) * 8 + rq] : 0.0;
#pragma unroll
        for (int I = 0; I < NTI; ++I)
          if (I >= J) T[tidx(I, J)] = __builtin_amdgcn_mfma_f64_16x16x4f64(aA[I], bJ, T[tidx(I, J)], 0, 0, 0);
      }
      WSYNC();
    }
    // the column's 16 pivots on the tiles right of it, the next column's tiles first
    if (J + 1 < NTI) {
      double Lx[NTI][4], pv[4];
#pragma unroll
      for (int kq = 0; kq < 4; ++kq) {
        const int cc = 16 * J + 4 * kq + rq, ccl = cc < m ? cc : m - 1;
        pv[kq] = cc < m ? c.Lp[ccl * (ccl + 3) / 2] : 0.0;
      }
#pragma unroll
      for (int X = 0; X < NTI; ++X) {
        const int row = 16 * X + cq, rr = row < m ? row : m - 1;
#pragma unroll
        for (int kq = 0; kq < 4; ++kq)
          Lx[X][kq] = (X > J && row < m) ? c.Lp[rr * (rr + 1) / 2 + 16 * J + 4 * kq + rq] : 0.0;
      }
#pragma unroll
      for (int Jp = 0; Jp < NTI; ++Jp) {
        if (Jp <= J) continue;
#pragma unroll
        for (int I = 0; I < NTI; ++I) {
          if (I < Jp) continue;
#pragma unroll
          for (int kq = 0; kq < 4; ++kq)
            T[tidx(I, Jp)] = __builtin_amdgcn_mfma_f64_16x16x4f64(-Lx[I][kq], Lx[Jp][kq] * pv[kq], T[tidx(I, Jp)], 0,
                                                                  0, 0);
        }
      }
      WSYNC();
    }
  }
  return nrm;
}

// The same factorisation and residuals on the upper triangle of K, the form the uniform batches
// use.  Tile (J, I), J <= I, holds element (16J + 4r + l/16, 16I + l%16) in register r of lane l,
// so the 4 pivot rows of a panel are one register of each tile of the panel's tile row, and that
// register is already the B operand (and, scaled by 1/pivot per lane group, the A operand) of
// v_mfma_f64_16x16x4 for the panel's rank-4 update of every trailing tile: no data movement
// between the elimination and the matrix cores.  Inside a panel every lane eliminates the 4 x 4
// diagonal block (read with lane broadcasts) for itself, and the pivot rows follow as
// U = L_block^{-1} R from lane-group broadcasts of the original rows, off the pivot chain.  The trailing updates act on whole tiles; entries of
// already eliminated rows and columns only ever receive rounding-level values and are never read.
template <bool RES, typename TS, int NMAX, int MIC>
__device__ __forceinline__ double factor_up(const Ctx<TS, NMAX, MIC>& c, St<TS, NMAX, MIC>& st, bool unit_d,
                                            double kappa) {
  constexpr int NTI = NMAX / 16;
  constexpr int NTILE = NTI * (NTI + 1) / 2;
  const int n = opaque_s(c.n), np = opaque_s(c.np), mi = opaque_s(c.mi), m = opaque_s(c.m), l = opaque_v(c.l);
  const int md = opaque_s(st.md);
  const int rq = l >> 4, cq = l & 15;
  double* vz = c.vb + mi;  // z, for G^T z
  double* kx = c.col + 64; // K's constant part times [x; y], then plus G^T z
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = 64 * cc + l;
    st.D[cc] = unit_d ? 1.0 : st.z[cc] / st.s[cc];
    if (k < mi) {
      c.vb[k] = st.D[cc];
      if (RES) vz[k] = st.z[cc];
    }
  }
  if (RES) c.col[l] = st.xa;
  // tiles start as K's constant part, loaded by the previous factorisation (or the problem load)
  d4 T[NTILE];
#pragma unroll
  for (int t = 0; t < NTILE; ++t) T[t] = st.Tk[t];
  WSYNC();
  double xc[NTI], pc[NTI], zacc[NTI];
#pragma unroll
  for (int I = 0; I < NTI; ++I) {
    xc[I] = RES ? c.col[16 * I + cq] : 0.0;
    pc[I] = 0.0;
    zacc[I] = 0.0;
  }
  if (RES) {
    // K's constant part times [x; y]: the stored upper tiles act directly on their rows and, off
    // the diagonal, transposed on their columns
    double pr[NTI][4];
#pragma unroll
    for (int J = 0; J < NTI; ++J) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double xr = c.col[16 * J + 4 * r + rq];
        double a = 0.0;
#pragma unroll
        for (int I = J; I < NTI; ++I) {
          a += T[tidx(I, J)][r] * xc[I];
          if (I > J) pc[I] += T[tidx(I, J)][r] * xr;
        }
        pr[J][r] = a;
      }
    }
#pragma unroll
    for (int J = 0; J < NTI; ++J) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double v = row16_sum(pr[J][r]);
        if (cq == 0) kx[16 * J + 4 * r + rq] = v;
      }
    }
  }
  // G^T D G into the tiles (operands as in factor, A from the tile row, B from the tile column),
  // with G x and G^T z under the MFMAs when RES.  When G reaches only 1..4 columns of the last tile
  // column (n = 50 in a 64-wide K: columns 48, 49), that tile column takes its products from
  // 4x4x4 MFMAs over a 4-column strip, at a quarter of the cost of the 16x16x4 tiles.  In the 4x4x4
  // layout, block b (lanes with (l & 15) >> 2 == b) holds A[i][k] in lane 16k + 4b + i, B[k][j] in
  // lane 16k + 4b + j and D[i][j] in lane 16i + 4b + j: the 16x16x4 A operand of a tile row is
  // then the A operand of its 16 rows in four blocks, and every block takes the strip's B.
  {
    int off[NTI];
    double msk[NTI];
#pragma unroll
    for (int I = 0; I < NTI; ++I) {
      const int cl = 16 * I + cq;
      off[I] = cl < n ? cl : n - 1;
      msk[I] = cl < n ? 1.0 : 0.0;
    }
    const int rem = n - 16 * (NTI - 1);
    const bool strip = NTI > 1 && rem >= 1 && rem <= 4;
    const int cs = 16 * (NTI - 1) + (cq & 3);
    const int offs = cs < n ? cs : n - 1;
    const double msks = cs < n ? 1.0 : 0.0;
    // raw loads of a step's G row entries (input precision), D and z; converted later
    auto load_raw = [&](int k0, TS (&g)[NTI], TS& gs, double& dk, double& zk, bool with_strip) {
      const int kc = k0 + rq < md ? k0 + rq : (md > 0 ? md - 1 : 0);
      dk = c.vb[kc];
      zk = RES ? vz[kc] : 0.0;
      const TS* grow = c.Gs + kc * n;
#pragma unroll
      for (int I = 0; I < NTI; ++I) g[I] = grow[off[I]];
      gs = with_strip ? grow[offs] : (TS)0;
    };
    auto gtdg = [&](auto tag) {
      constexpr bool S = decltype(tag)::value;
      double gc[NTI], zc, ac[NTI], gsc = 0.0, sacc[NTI];
#pragma unroll
      for (int J = 0; J < NTI; ++J) sacc[J] = 0.0;
      {
        TS g0[NTI], gs0;
        double d0, z0;
        load_raw(0, g0, gs0, d0, z0, S);
        const bool live = rq < md;
#pragma unroll
        for (int I = 0; I < NTI; ++I) {
          gc[I] = (double)g0[I] * msk[I];
          ac[I] = gc[I] * (live ? d0 : 0.0);
        }
        if (S) gsc = (double)gs0 * msks;
        zc = live ? z0 : 0.0;
      }
      for (int k0 = 0; k0 < md; k0 += 4) {
        const int kn = k0 + 4 < md ? k0 + 4 : k0;
        TS gr[NTI], gsr;
        double dr, zr;
        load_raw(kn, gr, gsr, dr, zr, S);
        const bool live_n = kn + rq < md;
        double gx = 0.0;
        // the step's vector work, issued after its MFMAs (float64 MFMAs occupy the SIMD's vector
        // datapath, so nothing is hidden by interleaving them): this step's G x row sums and G^T z
        // share, then the next step's operands
        auto task = [&](int g) {
          if (RES) {
            if (g == 0) {
#pragma unroll
              for (int I = 0; I < NTI; ++I) gx += gc[I] * xc[I];
            }
            if (g == 1) gx += dpp<0xB1>(gx);
            if (g == 2) gx += dpp<0x4E>(gx);
            if (g == 3) gx += dpp<0x141>(gx);
            if (g == 4) {
              gx += dpp<0x140>(gx);
              const int k = k0 + rq;
              if (cq == 0 && k < md) c.vb[k] = gx;
            }
            if (g == 5) {
#pragma unroll
              for (int I = 0; I < NTI; ++I) zacc[I] += gc[I] * zc;
            }
          }
          if (g == 6) {
#pragma unroll
            for (int I = 0; I < NTI; ++I) gc[I] = (double)gr[I] * msk[I];
            if (S) gsc = (double)gsr * msks;
          }
          if (g == 7) {
            const double dn = live_n ? dr : 0.0;
#pragma unroll
            for (int I = 0; I < NTI; ++I) ac[I] = gc[I] * dn;
            zc = live_n ? zr : 0.0;
          }
        };
        // the MFMAs read this step's operands before the tasks replace them: keep copies
        double am[NTI], bm[NTI];
#pragma unroll
        for (int I = 0; I < NTI; ++I) {
          am[I] = ac[I];
          bm[I] = gc[I];
        }
        const double bs = gsc;
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int J = 0; J < NTI; ++J) {
#pragma unroll
          for (int I = J; I < NTI; ++I) {
            if (S && I == NTI - 1)
              sacc[J] = __builtin_amdgcn_mfma_f64_4x4x4f64(am[J], bs, sacc[J], 0, 0, 0);
            else
              T[tidx(I, J)] = __builtin_amdgcn_mfma_f64_16x16x4f64(am[J], bm[I], T[tidx(I, J)], 0, 0, 0);
          }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int g = 0; g < (NTILE > 8 ? NTILE : 8); ++g) task(g);
      }
      if (S) {
        // strip rows 16J + 4b + i, columns 16(NTI-1) + j (lane 16i + 4b + j) into register b of
        // tile (J, NTI-1), lane 16i + j
#pragma unroll
        for (int J = 0; J < NTI; ++J) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const double v = r == 0 ? sacc[J] : __shfl(sacc[J], l + 4 * r, 64);
            T[tidx(NTI - 1, J)][r] += cq < 4 ? v : 0.0;
          }
        }
      }
    };
    // (the strip's extra live registers cost more than it saves in the factorisations without the
    // fused residuals: the backward kernel's and the first one)
    if constexpr (RES) {
      if (strip)
        gtdg(BoolTag<true>{});
      else
        gtdg(BoolTag<false>{});
    } else {
      gtdg(BoolTag<false>{});
    }
  }
  WSYNC();
  const double zbox = box_rows_into(c, st, T, md, RES);
  double nrm = 0.0;
  if (RES) {
#pragma unroll
    for (int I = 0; I < NTI; ++I) {
      double v = pc[I] + zacc[I];
      v += __shfl_xor(v, 16, 64);
      v += __shfl_xor(v, 32, 64);
      pc[I] = v;
    }
    if (rq == 0) {
#pragma unroll
      for (int I = 0; I < NTI; ++I) kx[16 * I + cq] += pc[I];
    }
    WSYNC();
    st.ra = l < m ? kx[l] + st.qa + zbox : 0.0;
    nrm = fabs(st.ra);
#pragma unroll
    for (int cc = 0; cc < MIC; ++cc) {
      const int k = 64 * cc + l;
      if (k < mi) {
        st.r3[cc] = c.vb[k] + st.s[cc] - st.h[cc];
        st.r2[cc] = st.s[cc] * st.z[cc] - kappa;
        nrm = nmax(nrm, nmax(fabs(st.r2[cc]), fabs(st.r3[cc])));
      } else {
        st.r3[cc] = 0.0;
        st.r2[cc] = 0.0;
      }
    }
    WSYNC();
    nrm = wmax(nrm);
  }
  double* pb = c.pb;
  // reference diagonal of this lane's row: the entry before elimination; for Schur rows it is
  // replaced by the entry once all H pivots are done (the tiles are current at that panel)
  auto snap_diag = [&]() {
#pragma unroll
    for (int J = 0; J < NTI; ++J) {
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (4 * r + rq == cq) pb[16 * J + cq] = T[tidx(J, J)][r];
    }
    WSYNC();
    const double v = pb[l];
    WSYNC();
    return v;
  };
  double dref = snap_diag();
  // this lane's packed-factor row in each tile row I (row 16I + cq, column rq), addressed once
  // per factorisation; a panel's stores add its first column c0 as an immediate offset
  double* lrow[NTI];
  bool lin[NTI];
#pragma unroll
  for (int I = 0; I < NTI; ++I) {
    const int row = 16 * I + cq;
    lrow[I] = c.Lp + opaque_v(row * (row + 1) / 2 + rq);
    lin[I] = row < m;
  }
#pragma unroll
  for (int J = 0; J < NTI; ++J) {
    if (16 * J >= m) break;
#pragma unroll
    for (int pp = 0; pp < 4; ++pp) {
      const int c0 = 16 * J + 4 * pp, jc = 4 * pp;
      if (c0 == np) {
        const double v = snap_diag();
        if (l >= np) dref = v;
      }
      const double rf = kPivotRel * fabs(dref);
      double thr[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) thr[t] = rl(rf, c0 + t);
      // the panel's pivot rows: register pp of the tiles of tile row J (row c0 + rq in lane group rq)
      double P[NTI];
#pragma unroll
      for (int I = 0; I < NTI; ++I) P[I] = I >= J ? T[tidx(I, J)][pp] : 0.0;
      // the upper triangle of the panel's 4 x 4 diagonal block as uniform values (the block is
      // symmetric), and the broadcasts of its first three rows to every lane group (both depend on
      // the panel's input only)
      double V[4][4];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int u = i; u < 4; ++u) V[i][u] = rl(P[J], 16 * i + jc + u);
      double R[3][NTI];
#pragma unroll
      for (int t = 0; t < 3; ++t)
#pragma unroll
        for (int I = 0; I < NTI; ++I) R[t][I] = I >= J ? __shfl(P[I], 16 * t + cq, 64) : 0.0;
      // the broadcasts go out before the pivot chain, whose latency then covers theirs
      __builtin_amdgcn_sched_barrier(0);
      // eliminate the block (L D L^T on its upper triangle): pivots, and multipliers
      // lm[u][t] = U[t][u] / D[t] of the unit-lower block factor.  The chain from one pivot to the
      // next is kept short: the reciprocal is formed before the pivot test (a failed test selects
      // 1 / (+-kHugePivot) afterwards), and the next pivot is one fused operation on U[t][t+1]^2,
      // formed before 1 / pivot t is known.
      double inv[4], lm[4][4];
      double dj = V[0][0];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int cc = c0 + t;
        const double sgn = cc < np ? 1.0 : -1.0;
        const bool ok = sgn * dj > thr[t];
        const double r = recip_c(dj);
        inv[t] = ok ? r : sgn * (1.0 / kHugePivot);
        st.rdpiv = l == cc ? inv[t] : st.rdpiv;
        if (t == 3) break;
        dj = fma(-(V[t][t + 1] * V[t][t + 1]), inv[t], V[t + 1][t + 1]);
#pragma unroll
        for (int u = t + 1; u < 4; ++u) lm[u][t] = V[t][u] * inv[t];
#pragma unroll
        for (int u = t + 1; u < 4; ++u)
#pragma unroll
          for (int w = u; w < 4; ++w)
            if (w != t + 1) V[u][w] -= lm[u][t] * V[t][w];
      }
      // pivot rows U = L_block^{-1} R: row u gains sum_{t<u} Linv[u][t] * (original row t)
      const double i10 = -lm[1][0], i21 = -lm[2][1], i32 = -lm[3][2];
      const double i20 = -lm[2][0] - lm[2][1] * i10;
      const double i31 = -lm[3][1] - lm[3][2] * i21;
      const double i30 = -lm[3][0] - lm[3][1] * i10 - lm[3][2] * i20;
      // per lane group rq, selected on the bits of rq (selects, not branches)
      const bool b0 = rq & 1, b1 = rq & 2;
      auto sel4 = [&](double v0, double v1, double v2, double v3) {
        const double lo = b0 ? v1 : v0, hi = b0 ? v3 : v2;
        return b1 ? hi : lo;
      };
      const double k0 = sel4(0.0, i10, i20, i30);
      const double k1 = sel4(0.0, 0.0, i21, i31);
      const double k2 = b0 && b1 ? i32 : 0.0;
#pragma unroll
      for (int I = 0; I < NTI; ++I)
        if (I >= J) P[I] = fma(k0, R[0][I], fma(k1, R[1][I], fma(k2, R[2][I], P[I])));
      // this lane group's pivot reciprocal; multipliers L[16I + cq][c0 + rq] into the packed factor
      const double myinv = sel4(inv[0], inv[1], inv[2], inv[3]);
#pragma unroll
      for (int I = 0; I < NTI; ++I) {
        if (I < J) continue;
        const int row = 16 * I + cq;
        if ((I > J || row > c0 + rq) && lin[I]) lrow[I][c0] = P[I] * myinv;
      }
      // rank-4 update of every trailing tile: A = -L rows (P of the tile row, scaled), B = U rows
      // (P of the tile column); the panel's own tile row first, which the next panel reads.  The
      // last panel of a tile row leaves that row's tiles alone.
#pragma unroll
      for (int Jp = 0; Jp < NTI; ++Jp) {
        if (Jp < J || (Jp == J && pp == 3)) continue;
        const double a = -P[Jp] * myinv;
#pragma unroll
        for (int I = 0; I < NTI; ++I)
          if (I >= Jp) T[tidx(I, Jp)] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, P[I], T[tidx(I, Jp)], 0, 0, 0);
      }
    }
  }
  WSYNC();
  return nrm;
}

// Solve K [dx; dy] = [p1; v4] (qpax/pdip.py:49-63) with v1 = sg*a1 (augmented with v4 = sg*a4 in
// rows np..m-1: the argument pa), v2 = sg*a2, v3 = sg*a3.  Writes st.da (dx; 0; dy) and the given
// ds/dz chunks.
template <bool FAST, typename TS, int NMAX, int MIC>
__device__ __forceinline__ void solve(const Ctx<TS, NMAX, MIC>& c, St<TS, NMAX, MIC>& st, double pa,
                                      const double (&a2)[MIC], const double (&a3)[MIC], double sg,
                                      double (&ods)[MIC], double (&odz)[MIC]) {
  const int n = opaque_s(c.n), mi = opaque_s(c.mi), m = opaque_s(c.m), l = opaque_v(c.l), md = opaque_s(st.md);
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = 64 * cc + l;
    if (k < mi) c.vb[k] = st.D[cc] * (sg * a3[cc] - (sg * a2[cc]) / st.z[cc]);
  }
  WSYNC();
  if (l < n) pa += gcol_dot(c.Gs, n, md, l, c.vb) + (md < mi ? box_gather<1>(c.Gs, n, md, mi, l, c.vb) : 0.0);
  if (l >= m) pa = 0.0;
  // forward: unit-lower L y = p; lane l reads row lr = min(l, m-1) of L.  The factor entries
  // load 16 at a time ahead of the dependency chain.  Reads past the end of a row run into the
  // following rows or the panel buffer behind the factor (inside this wavefront's LDS image) and
  // are discarded by the select, which keeps only columns j < lr: the factorisation writes no
  // diagonal slot, so lanes of rows >= m (which read row m-1) must not take L[m-1][m-1] either --
  // their values stay finite whatever the LDS held before, and lanes j >= m only ever meet zero
  // coefficients.
  constexpr int NB = NMAX / 16;
  const int lr = l < m ? l : m - 1;
  const double* Lrow = c.Lp + lr * (lr + 1) / 2;
  {
    double lv[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) lv[u] = Lrow[u];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      double nv[16];
      if (b + 1 < NB) {
#pragma unroll
        for (int u = 0; u < 16; ++u) nv[u] = Lrow[16 * (b + 1) + u];
      }
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int j = 16 * b + u;
        pa -= (lr > j ? lv[u] : 0.0) * rl(pa, j);
      }
      if (b + 1 < NB) {
#pragma unroll
        for (int u = 0; u < 16; ++u) lv[u] = nv[u];
      }
    }
  }
  pa = l < m ? pa * st.rdpiv : 0.0;
  // backward: L^T x = y, column j of L^T is row j of L (rows clamped to m - 1 for the address).
  // FAST: when the reads of rows m..NMAX-1 stay inside the panel buffer behind the factor (at
  // least 512 doubles), every block reads its rows at compile-time offsets, branch-free, with the
  // next block's loads issued a block ahead; the select drops what those rows hold.  Rows below
  // kMinTail are below every m that passes the test and skip the j < m term.
  constexpr int kTailReach = (NMAX - 1) * NMAX / 2 + 64 - 512;
  constexpr int kMinTail = [] {
    int mm = 1;
    while (mm * (mm + 1) / 2 < kTailReach) ++mm;
    return mm;
  }();
  if (FAST && m * (m + 1) / 2 >= kTailReach) {
    double lv[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) lv[u] = c.Lp[(NMAX - 1 - u) * (NMAX - u) / 2 + l];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      double nv[16];
      if (b + 1 < NB) {
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const int j = NMAX - 1 - 16 * (b + 1) - u;
          nv[u] = c.Lp[j * (j + 1) / 2 + l];
        }
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int j = NMAX - 1 - 16 * b - u;
        if (j > 0) pa -= ((j < kMinTail || j < m) && l < j ? lv[u] : 0.0) * rl(pa, j);
      }
      if (b + 1 < NB) {
#pragma unroll
        for (int u = 0; u < 16; ++u) lv[u] = nv[u];
      }
    }
  } else {
    double lv[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int j = NMAX - 1 - u, jj = j < m ? j : m - 1;
      lv[u] = c.Lp[jj * (jj + 1) / 2 + l];
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      double nv[16];
      if (b + 1 < NB) {
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const int j = NMAX - 1 - (16 * (b + 1) + u), jj = j < m ? j : m - 1;
          nv[u] = c.Lp[jj * (jj + 1) / 2 + l];
        }
      }
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int j = NMAX - 1 - (16 * b + u);
        if (j > 0) pa -= (l < j && j < m ? lv[u] : 0.0) * rl(pa, j);
      }
      if (b + 1 < NB) {
#pragma unroll
        for (int u = 0; u < 16; ++u) lv[u] = nv[u];
      }
    }
  }
  st.da = l < m ? pa : 0.0;
  c.col[l] = st.da;
  WSYNC();
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = 64 * cc + l;
    if constexpr (FAST) {
      // every lane forms a row product (rows past the dense ones clamped to the last of them) and
      // selects; box rows take their single entry instead
      const int kc = k < md ? k : (md > 0 ? md - 1 : 0);
      double g = (n & 1) == 0 ? grow_dot2(c.Gs, n, kc, c.col) : grow_dot(c.Gs, n, kc, c.col);
      if (md < mi) {
        const double gb = box_row(c.Gs, n, k < mi ? k : mi - 1, c.col);
        g = k < md ? g : gb;
      }
      const double v3 = sg * a3[cc], v2 = sg * a2[cc];
      const double dsk = v3 - g;
      const double dzk = (v2 - st.z[cc] * dsk) / st.s[cc];
      ods[cc] = k < mi ? dsk : 0.0;
      odz[cc] = k < mi ? dzk : 0.0;
    } else {
      if (k < mi) {
        const double g = k < md ? grow_dot(c.Gs, n, k, c.col) : box_row(c.Gs, n, k, c.col);
        const double v3 = sg * a3[cc], v2 = sg * a2[cc];
        const double dsk = v3 - g;
        ods[cc] = dsk;
        odz[cc] = (v2 - st.z[cc] * dsk) / st.s[cc];
      } else {
        ods[cc] = 0.0;
        odz[cc] = 0.0;
      }
    }
  }
  WSYNC();
}

// FAST: the ratio is formed in every lane and selected (no divergent branch around the division).
template <bool FAST, int MIC>
__device__ __forceinline__ double ort(const double (&x)[MIC], const double (&dx)[MIC], int mi, int l) {
  double a = 1.0;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = 64 * cc + l;
    if constexpr (FAST) {
      const double r = -x[cc] / dx[cc];
      a = (k < mi && dx[cc] < 0) ? nmin(a, r) : a;
    } else {
      if (k < mi && dx[cc] < 0) a = nmin(a, -x[cc] / dx[cc]);
    }
  }
  return wmin(a);
}

template <bool FAST, typename TS, int NMAX, int MIC>
__device__ __forceinline__ void init(const Ctx<TS, NMAX, MIC>& c, St<TS, NMAX, MIC>& st) {
  const int mi = c.mi, l = c.l;
  if constexpr (FAST) factor_up<false>(c, st, true, 0.0);
  else factor<false>(c, st, true, 0.0);
  // p = [-q + G^T h; b]: with a2 = 0, a3 = h, sg = 1 the solve forms p1 = v1 + G^T h
  double zero[MIC];
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) zero[cc] = 0.0;
  const double pa = -st.qa;  // -q in rows < n, b in rows np..m-1 (qa holds -b), 0 elsewhere
  if constexpr (FAST) tile_load(c, st.Tk);  // the next factorisation's tiles, behind this solve
  solve<FAST>(c, st, pa, zero, st.h, 1.0, st.ds, st.dz);
  st.xa = st.da;
  // z = G x - h  ==  -(h - G x) = -ds from the solve above
  double zmax = -INFINITY, zmin = INFINITY;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = 64 * cc + l;
    const double zk = -st.ds[cc];
    st.z[cc] = zk;
    if (k < mi) { zmax = nmax(zmax, zk); zmin = nmin(zmin, zk); }
  }
  zmax = wmax(zmax);
  zmin = wmin(zmin);
  const double ap = zmax, ad = -zmin;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const double zk = st.z[cc];
    st.s[cc] = (ap < 0) ? -zk : -zk + (1 + ap);
    st.z[cc] = (ad >= 0) ? zk + (1 + ad) : zk;
  }
}

template <bool FAST, typename TS, int NMAX, int MIC>
__device__ __forceinline__ void step(const Ctx<TS, NMAX, MIC>& c, St<TS, NMAX, MIC>& st, bool relax_rule) {
  const double as = ort<FAST, MIC>(st.s, st.ds, c.mi, c.l);
  const double az = ort<FAST, MIC>(st.z, st.dz, c.mi, c.l);
  const double alpha = relax_rule ? 0.99 * nmin(nmin(1.0, 0.99 * as), 0.99 * az) : 0.99 * nmin(as, az);
  st.xa += alpha * st.da;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    st.s[cc] += alpha * st.ds[cc];
    st.z[cc] += alpha * st.dz[cc];
  }
}

// One predictor-corrector iteration (qpax/pdip.py:87-119); returns the entry convergence flag.
template <bool FRES, typename TS, int NMAX, int MIC>
__device__ __forceinline__ int pc_iter(const Ctx<TS, NMAX, MIC>& c, St<TS, NMAX, MIC>& st, double tol) {
  const int mi = c.mi, l = c.l;
  double nrm;
  if constexpr (FRES) {
    nrm = factor_up<true>(c, st, false, 0.0);
  } else {
    nrm = residuals(c, st, 0.0);
    factor<false>(c, st, false, 0.0);
  }
  solve<FRES>(c, st, -st.ra, st.r2, st.r3, -1.0, st.dsa, st.dza);
  // centering (qpax/pdip.py:76-84)
  double sz = 0.0;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) if (64 * cc + l < mi) sz += st.s[cc] * st.z[cc];
  sz = wsum(sz);
  const double mu = sz / mi;
  const double a = nmin(ort<FRES, MIC>(st.s, st.dsa, mi, l), ort<FRES, MIC>(st.z, st.dza, mi, l));
  double gg = 0.0;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc)
    if (64 * cc + l < mi) gg += (st.s[cc] + a * st.dsa[cc]) * (st.z[cc] + a * st.dza[cc]);
  gg = wsum(gg);
  const double ratio = gg / sz;
  const double sm = ratio * ratio * ratio * mu;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) st.r2[cc] = st.r2[cc] - (sm - st.dsa[cc] * st.dza[cc]);
  if constexpr (FRES) tile_load(c, st.Tk);  // the next factorisation's tiles, behind this solve
  solve<FRES>(c, st, -st.ra, st.r2, st.r3, -1.0, st.ds, st.dz);
  step<FRES>(c, st, false);
  return nrm < tol ? 1 : 0;
}

// One relaxed Newton iteration (qpax/pdip_relaxed.py:7-52).
template <bool FRES, typename TS, int NMAX, int MIC>
__device__ __forceinline__ int relax_iter(const Ctx<TS, NMAX, MIC>& c, St<TS, NMAX, MIC>& st, double tol,
                                          double kappa) {
  double nrm;
  if constexpr (FRES) {
    nrm = factor_up<true>(c, st, false, kappa);
  } else {
    nrm = residuals(c, st, kappa);
    factor<false>(c, st, false, 0.0);
  }
  if constexpr (FRES) tile_load(c, st.Tk);  // the next factorisation's tiles, behind this solve
  solve<FRES>(c, st, -st.ra, st.r2, st.r3, -1.0, st.ds, st.dz);
  step<FRES>(c, st, true);
  return nrm < tol ? 1 : 0;
}

// Context of one problem for the calling wavefront; several wavefronts of a workgroup each own
// a disjoint LDS image of lds_per_wave bytes.
template <typename TS, int NMAX, int MIC>
__device__ __forceinline__ Ctx<TS, NMAX, MIC> make_ctx(int n, int me, int mi, double* ws, int lds_per_wave) {
  QPX_DYN_LDS(qpx_smem);
  const Lds w = lds_layout(n, me, mi, (int)sizeof(TS));
  unsigned char* base = qpx_smem + (threadIdx.x >> 6) * lds_per_wave;
  Ctx<TS, NMAX, MIC> c;
  c.n = n; c.np = padded_n(n); c.me = me; c.mi = mi; c.m = c.np + me; c.l = threadIdx.x & 63;
  c.Gs = reinterpret_cast<const TS*>(base + w.off_G);
  c.Lp = reinterpret_cast<double*>(base + w.off_L);
  c.pb = reinterpret_cast<double*>(base + w.off_P);
  c.col = c.pb;
  c.vb = c.pb + 128;
  c.KQ = ws;
  return c;
}

template <typename TS, int NMAX, int MIC>
__device__ __forceinline__ void store_state(const Ctx<TS, NMAX, MIC>& c, const St<TS, NMAX, MIC>& st, const PIdx& pi,
                                            TS* x, TS* s, TS* z, TS* y, bool post) {
  const int n = c.n, np = c.np, me = c.me, mi = c.mi, l = c.l;
  if (l < n) x[pi.oq + l] = (TS)st.xa;
  else if (l >= np && l < np + me) y[pi.ob + (l - np)] = (TS)st.xa;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = 64 * cc + l;
    if (k < mi) {
      double sv = st.s[cc], zv = st.z[cc];
      if (post && st.hinf[cc]) { zv = 0.0; sv = INFINITY; }
      s[pi.oh + k] = (TS)sv;
      z[pi.oh + k] = (TS)zv;
    }
  }
}

// Next problem for the calling wavefront.  Without a work queue wavefront w of workgroup b takes
// entry b * (waves per workgroup) + w of the schedule once; with one, lane 0 pulls the next entry.
template <typename TS>
__device__ __forceinline__ int64_t next_problem(const Args<TS>& a, int& round) {
  int64_t idx;
  if (a.qhead) {
    unsigned v = 0;
    if ((threadIdx.x & 63) == 0) v = atomicAdd(a.qhead, 1u);
    idx = (int64_t)__builtin_amdgcn_readfirstlane(v);
  } else {
    if (round++ > 0) return -1;
    idx = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  }
  const int64_t count = a.order ? a.order_count : a.B;
  if (idx >= count) return -1;
  return a.order ? (int64_t)a.order[idx] : idx;
}

// Forward kernel: solve (and optionally relax) QPs, one per wavefront at a time.  FRES selects
// factor_up with the residuals formed inside its G^T D G pass (uniform batches) over the separate
// residual pass and factor (ragged batches).
// Forward launch of a uniform batch whose last problems run their relaxations as work items of
// their own: tickets 0..F-1 solve and relax problems 0..F-1 in one item, F..B-1 are the PC
// solves of problems F..B-1, and B..2B-F-1 their relaxations.  A split solve leaves its
// post-processed float64 state (s = inf, z = 0 on rows with infinite h) after the constant
// matrix in its workspace slice and raises a ready flag (release); the relaxation of the same
// problem waits for the flag (acquire), reloads G into LDS and continues from that state.  A
// relaxation's ticket is taken after its solve's, which a running wavefront holds, so the wait
// ends; the wait is bounded all the same (after a.split_spin polls, kSplitSpin unless the host
// lowers it, the relaxation writes NaN to its relaxed point, reports riters = -1 and stops).  The
// last work items of the launch are then relaxations (about 4 iterations) instead of whole
// problems, and only the last problems pay the second load.
constexpr int kSplitSpin = 1 << 22;

template <typename TS, int NMAX, int MIC>
__device__ __forceinline__ void wave_split(const Args<TS>& a, int64_t F) {
  unsigned tk = 0;
  if ((threadIdx.x & 63) == 0) tk = atomicAdd(a.qhead, 1u);
  const int64_t u = (int64_t)__builtin_amdgcn_readfirstlane(tk);
  if (u >= 2 * a.B - F) return;
  const bool rel = u >= a.B, fused = u < F;
  const int64_t p = rel ? u - a.B + F : u;
  const PIdx pi = pidx(a, p);
  const auto c = make_ctx<TS, NMAX, MIC>(pi.n, pi.me, pi.mi, a.ws + pi.ows, a.lds_per_wave);
  double* sw = a.ws + pi.ows + split_state_off(pi.n, pi.me);
  unsigned* ready = reinterpret_cast<unsigned*>(sw + kSplitStateWords);
  const int l = c.l;
  if (rel) {
    bool ok = false;
    for (int t = 0; t < a.split_spin; ++t) {
      if (__hip_atomic_load(ready, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
        ok = true;
        break;
      }
      __builtin_amdgcn_s_sleep(8);
    }
    if (!ok) {
      // no relaxed point: NaN, so that nothing downstream (the backward) takes garbage for one
      const TS nan = (TS)NAN;
      if (l < pi.n) a.xr[pi.oq + l] = nan;
      if (l < pi.me) a.yr[pi.ob + l] = nan;
      for (int k = l; k < pi.mi; k += 64) {
        a.sr[pi.oh + k] = nan;
        a.zr[pi.oh + k] = nan;
      }
      if (l == 0) { a.rconv[p] = 0; a.riters[p] = -1; }
      return;
    }
  }
  St<TS, NMAX, MIC> st;
  load<true>(c, st, a.Q + pi.oQ, a.q + pi.oq, a.A + pi.oA, a.b + pi.ob, a.G + pi.oG, a.h + pi.oh, !rel);
  if (!rel) {
    init<true>(c, st);
    int conv = 0, it = 0;
    while (it < kMaxIters && conv == 0) {
      conv = pc_iter<true>(c, st, a.tol);
      ++it;
    }
    store_state(c, st, pi, a.x, a.s, a.z, a.y, true);
    if (l == 0) { a.conv[p] = conv; a.iters[p] = it; }
    if (!fused) {
      sw[l] = st.xa;
#pragma unroll
      for (int cc = 0; cc < MIC; ++cc) {
        sw[64 + 128 * cc + l] = st.hinf[cc] ? INFINITY : st.s[cc];
        sw[128 + 128 * cc + l] = st.hinf[cc] ? 0.0 : st.z[cc];
      }
      __threadfence();
      if (l == 0) __hip_atomic_store(ready, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
    // relax from the post-processed point (s = inf, z = 0 on rows with infinite h)
#pragma unroll
    for (int cc = 0; cc < MIC; ++cc)
      if (st.hinf[cc]) { st.z[cc] = 0.0; st.s[cc] = INFINITY; }
  } else {
    st.xa = sw[l];
#pragma unroll
    for (int cc = 0; cc < MIC; ++cc) {
      st.s[cc] = sw[64 + 128 * cc + l];
      st.z[cc] = sw[128 + 128 * cc + l];
    }
  }
  {
    int rconv = 0, rit = 0;
    while (rit < kMaxIters && rconv == 0) {
      rconv = relax_iter<true>(c, st, a.tol, a.kappa);
      ++rit;
    }
    store_state(c, st, pi, a.xr, a.sr, a.zr, a.yr, false);
    if (l == 0) { a.rconv[p] = rconv; a.riters[p] = rit; }
  }
}

// Resets the ticket counter and the ready flags of a split forward launch (a kernel, so that a
// captured graph replays it).
template <typename TS>
__global__ __launch_bounds__(256) void k_wave_split_clear(Args<TS> a, int64_t flag_word) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0) a.qhead[0] = 0u;
  if (p < a.B) *reinterpret_cast<unsigned*>(a.ws + p * a.stride + flag_word) = 0u;
}

template <typename TS, int NMAX, int MIC, bool FRES>
__global__ __launch_bounds__(256) void k_wave_solve(Args<TS> a, int with_relax) {
  if constexpr (FRES) {
    if (with_relax == 2) {
      wave_split<TS, NMAX, MIC>(a, a.split_fused);
      return;
    }
  }
  int round = 0;
  for (int64_t p; (p = next_problem(a, round)) >= 0;) {
    const PIdx pi = pidx(a, p);
    const int n = pi.n, me = pi.me, mi = pi.mi;
    const auto c = make_ctx<TS, NMAX, MIC>(n, me, mi, a.ws + pi.ows, a.lds_per_wave);
    St<TS, NMAX, MIC> st;
    load<FRES>(c, st, a.Q + pi.oQ, a.q + pi.oq, a.A + pi.oA, a.b + pi.ob, a.G + pi.oG, a.h + pi.oh);
    init<FRES>(c, st);
    int conv = 0, it = 0;
    while (it < kMaxIters && conv == 0) {
      conv = pc_iter<FRES>(c, st, a.tol);
      ++it;
    }
    store_state(c, st, pi, a.x, a.s, a.z, a.y, true);
    if (c.l == 0) { a.conv[p] = conv; a.iters[p] = it; }
    if (!with_relax) continue;
    // relax from the post-processed point (s = inf, z = 0 on rows with infinite h)
#pragma unroll
    for (int cc = 0; cc < MIC; ++cc)
      if (st.hinf[cc]) { st.z[cc] = 0.0; st.s[cc] = INFINITY; }
    int rconv = 0, rit = 0;
    while (rit < kMaxIters && rconv == 0) {
      rconv = relax_iter<FRES>(c, st, a.tol, a.kappa);
      ++rit;
    }
    store_state(c, st, pi, a.xr, a.sr, a.zr, a.yr, false);
    if (c.l == 0) { a.rconv[p] = rconv; a.riters[p] = rit; }
    WSYNC();
  }
}

// Backward kernel (qpax/diff_qp.py:20-46 with optnet_derivatives :8-17); FAST selects factor_up
// (uniform batches) over factor (ragged batches).
template <typename TS, int NMAX, int MIC, bool FAST>
__global__ __launch_bounds__(256) void k_wave_bwd(Args<TS> a) {
  int round = 0;
  for (int64_t p; (p = next_problem(a, round)) >= 0;) {
    const PIdx pi = pidx(a, p);
    const int n = pi.n, me = pi.me, mi = pi.mi;
    const auto c = make_ctx<TS, NMAX, MIC>(n, me, mi, a.ws + pi.ows, a.lds_per_wave);
    const int l = c.l, np = c.np, m = c.m;
    St<TS, NMAX, MIC> st;
    load<FAST>(c, st, a.Q + pi.oQ, a.q + pi.oq, a.A + pi.oA, a.b + pi.ob, a.G + pi.oG, a.h + pi.oh);
    st.xa = l < n ? (double)a.xr_in[pi.oq + l] : ((l >= np && l < m) ? (double)a.yr_in[pi.ob + (l - np)] : 0.0);
#pragma unroll
    for (int cc = 0; cc < MIC; ++cc) {
      const int k = 64 * cc + l;
      st.s[cc] = k < mi ? (double)a.sr_in[pi.oh + k] : 1.0;
      st.z[cc] = k < mi ? (double)a.zr_in[pi.oh + k] : 1.0;
    }
    if constexpr (FAST) factor_up<false>(c, st, false, 0.0);
    else factor<false>(c, st, false, 0.0);
    double zero[MIC];
#pragma unroll
    for (int cc = 0; cc < MIC; ++cc) zero[cc] = 0.0;
    const double g = l < n ? (double)a.g[pi.oq + l] : 0.0;
    solve<FAST>(c, st, -g, zero, zero, 1.0, st.ds, st.dz);
    // cotangents; dz holds dlam_tilde, dlam = dlam_tilde / lam
    double* lamv = c.vb;        // lam = z
    double* dlamv = c.vb + mi;  // dlam = dlam_tilde / lam
    c.col[l] = st.xa;
    c.col[64 + l] = st.da;
#pragma unroll
    for (int cc = 0; cc < MIC; ++cc) {
      const int k = 64 * cc + l;
      if (k < mi) {
        lamv[k] = st.z[cc];
        dlamv[k] = st.dz[cc] / st.z[cc];
      }
    }
    WSYNC();
    const double* X = c.col;
    const double* DX = c.col + 64;
    // lane j owns column j (n <= 64); the per-row scalars are read eight rows at a time (uniform
    // LDS broadcasts issued ahead of the arithmetic), so no row waits on its own LDS round trip
    const bool cj = l < n;
    const int jl = cj ? l : 0;
    const double xj = X[jl], dxj = DX[jl];
    auto rows8 = [&](int nrow, const double* r1, const double* r2, auto&& emit) {
      int k0 = 0;
      for (; k0 + 8 <= nrow; k0 += 8) {
        double u1[8], u2[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          u1[u] = r1[k0 + u];
          u2[u] = r2[k0 + u];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) emit(k0 + u, u1[u], u2[u]);
      }
      for (; k0 < nrow; ++k0) emit(k0, r1[k0], r2[k0]);
    };
    rows8(n, DX, X, [&](int i, double dxi, double xi) {
      if (cj) a.dQ[pi.oQ + i * n + l] = (TS)(0.5 * (dxi * xj + xi * dxj));
    });
    if (cj) a.dq[pi.oq + l] = (TS)st.da;
    rows8(me, DX + np, X + np, [&](int cix, double dnu, double nu) {
      if (cj) a.dA[pi.oA + cix * n + l] = (TS)(dnu * xj + nu * dxj);
    });
    if (l >= np && l < m) a.db[pi.ob + (l - np)] = (TS)(-st.da);
    const bool exact = a.flags & 1;
    rows8(mi, lamv, dlamv, [&](int k, double lam, double dlam) {
      const double v = exact ? lam * dlam * xj + lam * dxj : lam * (dlam * xj + lam * dxj);
      if (cj) a.dG[pi.oG + k * n + l] = (TS)v;
    });
#pragma unroll
    for (int cc = 0; cc < MIC; ++cc) {
      const int k = 64 * cc + l;
      if (k < mi) a.dh[pi.oh + k] = (TS)(-st.z[cc] * (st.dz[cc] / st.z[cc]));
    }
    WSYNC();
  }
}

}  // namespace wave
}  // namespace qpx
