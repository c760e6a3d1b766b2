// Fused solver for small dense QPs: one wavefront owns one QP and keeps it on chip.
//
// Layout (lane l = 0..63 of the wavefront, m = n + m_eq <= NMAX <= 64):
//   * registers: row l of the augmented KKT matrix K = [[Q, A^T], [A, 0]] (kq, constant) and of
//     the working matrix / unit-lower factor (kr); the augmented vectors [x; y], [r1; r4],
//     [dx; dy] one element per lane; the m_ineq-vectors (s, z, D, residuals, directions) in
//     MIC chunks of 64 rows, row 64c + l in lane l.
//   * LDS: G and A (input precision), the packed lower factor (for the transposed solve), two
//     64-entry broadcast buffers and one m_ineq-entry broadcast buffer.
// The reduced KKT system of qpax's cvxopt reduction (qpax/pdip.py:38-63),
//     [[H, A^T], [A, 0]] [dx; dy] = [p1; v4],   H = Q + G^T diag(z/s) G,
// is factored as one quasi-definite LDL^T (pivots 0..n-1 positive: the Cholesky of H; pivots
// n..m-1 negative: minus the Schur complement F = A H^{-1} A^T), which is the same elimination
// as the reference's H-then-F factorisation in a single pass.  G^T D G runs on the float64
// matrix cores (v_mfma_f64_16x16x4); everything else on the vector ALUs.
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

__device__ __forceinline__ double wmax(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = nmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wmin(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = nmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <typename TS, int NMAX, int MIC>
struct St {
  // constant problem data (the augmented matrix itself lives in the workspace, see Ctx::KQ)
  double qa;      // q (rows < n), -b (rows n..m-1)
  double h[MIC];
  bool hinf[MIC];
  // factor
  double kr[NMAX];
  double rdpiv;
  // iterate and work vectors
  double xa, ra, da;
  double s[MIC], z[MIC], D[MIC], r2[MIC], r3[MIC], dsa[MIC], dza[MIC], ds[MIC], dz[MIC];
};

template <typename TS, int NMAX, int MIC>
struct Ctx {
  int n, me, mi, m, l;
  // K = [[Q_sym, A^T], [A, 0]] (m x m, float64, symmetric) in this problem's global workspace;
  // lane l reads column entries KQ[j*m + l] = K[j][l] = K[l][j], coalesced across the lanes.
  double* KQ;
  const TS* Gs;
  const TS* As;
  double* Lp;
  double* col;
  double* vb;
};

#define WSYNC() __syncthreads()

// sum_k G[k][col] * v[k] over k < mi (v broadcast from LDS); four rows per trip.
template <typename TS>
__device__ __forceinline__ double gcol_dot(const TS* __restrict__ Gs, int n, int mi, int col,
                                           const double* __restrict__ v) {
  double a0 = 0.0, a1 = 0.0;
  int k = 0;
  for (; k + 4 <= mi; k += 4) {
    const double2 v01 = *reinterpret_cast<const double2*>(v + k);
    const double2 v23 = *reinterpret_cast<const double2*>(v + k + 2);
    const double g0 = (double)Gs[k * n + col], g1 = (double)Gs[(k + 1) * n + col];
    const double g2 = (double)Gs[(k + 2) * n + col], g3 = (double)Gs[(k + 3) * n + col];
    a0 += g0 * v01.x;
    a1 += g1 * v01.y;
    a0 += g2 * v23.x;
    a1 += g3 * v23.y;
  }
  for (; k < mi; ++k) a0 += (double)Gs[k * n + col] * v[k];
  return a0 + a1;
}

// sum_j G[row][j] * v[j] over j < n (v broadcast from LDS); four columns per trip.
template <typename TS>
__device__ __forceinline__ double grow_dot(const TS* __restrict__ Gs, int n, int row, const double* __restrict__ v) {
  const TS* g = Gs + row * n;
  double a0 = 0.0, a1 = 0.0;
  int j = 0;
  for (; j + 4 <= n; j += 4) {
    const double2 v01 = *reinterpret_cast<const double2*>(v + j);
    const double2 v23 = *reinterpret_cast<const double2*>(v + j + 2);
    a0 += (double)g[j] * v01.x;
    a1 += (double)g[j + 1] * v01.y;
    a0 += (double)g[j + 2] * v23.x;
    a1 += (double)g[j + 3] * v23.y;
  }
  for (; j < n; ++j) a0 += (double)g[j] * v[j];
  return a0 + a1;
}

// 1/d to full double precision: hardware reciprocal plus two Newton steps.
__device__ __forceinline__ double recip(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  r = fma(r, fma(-d, r, 1.0), r);
  return r;
}

template <typename TS, int NMAX, int MIC>
__device__ __forceinline__ void load(const Ctx<TS, NMAX, MIC>& c, St<TS, NMAX, MIC>& st, const TS* Q,
                                     const TS* q, const TS* A, const TS* b, const TS* G, const TS* h) {
  const int n = c.n, me = c.me, mi = c.mi, m = c.m, l = c.l;
  TS* Gs = const_cast<TS*>(c.Gs);
  TS* As = const_cast<TS*>(c.As);
  for (int k = 0; k < mi; ++k) {
    const bool inf = isinf((double)h[k]);
    for (int j = l; j < n; j += 64) {
      const TS g = G[k * n + j];
      Gs[k * n + j] = inf ? (TS)0 * g : g;
    }
  }
  for (int e = l; e < me * n; e += 64) As[e] = A[e];
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = 64 * cc + l;
    const double hk = k < mi ? (double)h[k] : 0.0;
    st.hinf[cc] = isinf(hk);
    st.h[cc] = st.hinf[cc] ? 0.0 : hk;
  }
  if (l < m) {
    for (int j = 0; j < m; ++j) {
      double v = 0.0;
      if (l < n) {
        if (j < n) v = 0.5 * ((double)Q[l * n + j] + (double)Q[j * n + l]);
        else v = (double)A[(j - n) * n + l];
      } else if (j < n) {
        v = (double)A[(l - n) * n + j];
      }
      c.KQ[j * m + l] = v;
    }
  }
  __threadfence_block();
  st.qa = l < n ? (double)q[l] : (l < m ? -(double)b[l - n] : 0.0);
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) { st.s[cc] = 1.0; st.z[cc] = 1.0; }
  st.xa = 0.0;
  WSYNC();
}

// Residuals of the (relaxed) KKT system; returns the infinity norm (NaN-propagating).
template <typename TS, int NMAX, int MIC>
__device__ __forceinline__ double residuals(const Ctx<TS, NMAX, MIC>& c, St<TS, NMAX, MIC>& st, double kappa) {
  const int n = c.n, mi = c.mi, m = c.m, l = c.l;
  c.col[l] = st.xa;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = 64 * cc + l;
    if (k < mi) c.vb[k] = st.z[cc];
  }
  WSYNC();
  double acc = st.qa;
  if (l < m) {
    const double* kcol = c.KQ + l;
    double a1 = 0.0;
    int j = 0;
    for (; j + 2 <= m; j += 2) {
      const double2 x2 = *reinterpret_cast<const double2*>(c.col + j);
      acc += kcol[j * m] * x2.x;
      a1 += kcol[(j + 1) * m] * x2.y;
    }
    if (j < m) acc += kcol[j * m] * c.col[j];
    acc += a1;
  }
  if (l < n) acc += gcol_dot(c.Gs, n, mi, l, c.vb);
  st.ra = l < m ? acc : 0.0;
  double nrm = fabs(st.ra);
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = 64 * cc + l;
    if (k < mi) {
      const double g = grow_dot(c.Gs, n, k, c.col);
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

// D = z/s (or 1), K = kq + G^T D G on the top-left block, LDL^T of K in place.
template <typename TS, int NMAX, int MIC>
__device__ __forceinline__ void factor(const Ctx<TS, NMAX, MIC>& c, St<TS, NMAX, MIC>& st, bool unit_d) {
  constexpr int NTI = NMAX / 16;
  constexpr int NTILE = NTI * (NTI + 1) / 2;
  const int n = c.n, mi = c.mi, m = c.m, l = c.l;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = 64 * cc + l;
    st.D[cc] = unit_d ? 1.0 : st.z[cc] / st.s[cc];
    if (k < mi) c.vb[k] = st.D[cc];
  }
  WSYNC();
  d4 acc[NTILE];
#pragma unroll
  for (int t = 0; t < NTILE; ++t) acc[t] = d4{0.0, 0.0, 0.0, 0.0};
  const int nt = (n + 15) >> 4;
  const int kk = l >> 4, ii = l & 15;
  for (int k0 = 0; k0 < mi; k0 += 4) {
    const int k = k0 + kk;
    const bool kv = k < mi;
    const double dk = kv ? c.vb[k] : 0.0;
    double g[NTI];
#pragma unroll
    for (int I = 0; I < NTI; ++I) {
      const int cl = 16 * I + ii;
      g[I] = (kv && cl < n) ? (double)c.Gs[k * n + cl] : 0.0;
    }
    int t = 0;
#pragma unroll
    for (int I = 0; I < NTI; ++I) {
      const double a = g[I] * dk;
#pragma unroll
      for (int J = 0; J <= I; ++J, ++t) {
        if (I < nt) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, g[J], acc[t], 0, 0, 0);
      }
    }
  }
  // Move the tiles into row-per-lane layout through LDS (staging area = the factor's region).
  // Tile element (i, j) sits in lane j + 16*(i%4), register i/4.
#pragma unroll
  for (int j = 0; j < NMAX; ++j) st.kr[j] = (j < m && l < m) ? c.KQ[j * m + l] : 0.0;
  double diag = l < m ? c.KQ[l * m + l] : 0.0;
  {
    int t = 0;
#pragma unroll
    for (int I = 0; I < NTI; ++I) {
#pragma unroll
      for (int J = 0; J <= I; ++J, ++t) {
        if (I >= nt) continue;
        const int jj = l & 15, i0 = l >> 4;
#pragma unroll
        for (int r = 0; r < 4; ++r) c.Lp[(4 * r + i0) * 16 + jj] = acc[t][r];
        WSYNC();
        if ((l >> 4) == I) {
          const int row = l & 15;
#pragma unroll
          for (int q2 = 0; q2 < 16; q2 += 2) {
            const double2 v = *reinterpret_cast<const double2*>(c.Lp + row * 16 + q2);
            if (16 * I + row < n) {
              if (16 * J + q2 < n) st.kr[16 * J + q2] += v.x;
              if (16 * J + q2 + 1 < n) st.kr[16 * J + q2 + 1] += v.y;
              if (I == J) {
                if (q2 == row) diag += v.x;
                if (q2 + 1 == row) diag += v.y;
              }
            }
          }
        }
        WSYNC();
      }
    }
  }
  // Quasi-definite LDL^T, right-looking, one pivot per step; rows in lanes.
  double d0 = diag;
#pragma unroll
  for (int j = 0; j < NMAX; ++j) {
    if (j < m) {
      if (j == n && l >= n) d0 = diag;
      const double dj = rl(diag, j);
      const double d0j = rl(d0, j);
      const bool hp = j < n;
      const bool ok = hp ? (dj > kPivotRel * fabs(d0j)) : (dj < -kPivotRel * fabs(d0j));
      const double d = ok ? dj : (hp ? kHugePivot : -kHugePivot);
      const double inv = recip(d);
      if (l == j) st.rdpiv = inv;
      double* cb = c.col + (j & 1) * 64;
      const double cv = st.kr[j];
      cb[l] = cv;
      WSYNC();
      // rows above the pivot (and padding rows, which are zero) take no update
      const double lf = (l > j) ? cv * inv : 0.0;
      diag -= lf * cv;
      // columns k >= m hold zeros in every row, so the update runs to NMAX unguarded
#pragma unroll
      for (int k = (j + 1) & ~1; k < NMAX; k += 2) {
        const double2 c2 = *reinterpret_cast<const double2*>(cb + k);
        if (k > j) st.kr[k] -= lf * c2.x;
        st.kr[k + 1] -= lf * c2.y;
      }
      st.kr[j] = lf;
      if (l > j && l < m) c.Lp[l * (l + 1) / 2 + j] = lf;
    }
  }
  WSYNC();
}

// Solve K [dx; dy] = [p1; v4] for the cvxopt-reduced system of qpax/pdip.py:49-63 with
// v1 = sg*a1 (augmented with v4 = sg*a4 in lanes n..m-1), v2 = sg*a2, v3 = sg*a3.
// Writes st.da (dx; dy) and the given ds/dz chunks.
template <typename TS, int NMAX, int MIC>
__device__ __forceinline__ void solve(const Ctx<TS, NMAX, MIC>& c, St<TS, NMAX, MIC>& st, double pa,
                                      const double (&a2)[MIC], const double (&a3)[MIC], double sg,
                                      double (&ods)[MIC], double (&odz)[MIC]) {
  const int n = c.n, mi = c.mi, m = c.m, l = c.l;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = 64 * cc + l;
    if (k < mi) c.vb[k] = st.D[cc] * (sg * a3[cc] - (sg * a2[cc]) / st.z[cc]);
  }
  WSYNC();
  if (l < n) pa += gcol_dot(c.Gs, n, mi, l, c.vb);
  if (l >= m) pa = 0.0;
  // forward: unit-lower L y = p; kr[j] is zero in rows <= j
#pragma unroll
  for (int j = 0; j < NMAX; ++j) {
    if (j < m) pa -= st.kr[j] * rl(pa, j);
  }
  pa *= st.rdpiv;
  // backward: L^T x = y, rows of L from the packed copy in LDS.  The read for lanes >= j runs
  // past row j into the next rows / the broadcast buffers that follow the factor (in bounds of
  // this wavefront's LDS image) and is masked off.
  for (int j = m - 1; j > 0; --j) {
    const double lv = c.Lp[j * (j + 1) / 2 + l];
    pa -= (l < j ? lv : 0.0) * rl(pa, j);
  }
  st.da = l < m ? pa : 0.0;
  c.col[l] = st.da;
  WSYNC();
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = 64 * cc + l;
    if (k < mi) {
      const double g = grow_dot(c.Gs, n, k, c.col);
      const double v3 = sg * a3[cc], v2 = sg * a2[cc];
      const double dsk = v3 - g;
      ods[cc] = dsk;
      odz[cc] = (v2 - st.z[cc] * dsk) / st.s[cc];
    } else {
      ods[cc] = 0.0;
      odz[cc] = 0.0;
    }
  }
  WSYNC();
}

template <int MIC>
__device__ __forceinline__ double ort(const double (&x)[MIC], const double (&dx)[MIC], int mi, int l) {
  double a = 1.0;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = 64 * cc + l;
    if (k < mi && dx[cc] < 0) a = nmin(a, -x[cc] / dx[cc]);
  }
  return wmin(a);
}

template <typename TS, int NMAX, int MIC>
__device__ __forceinline__ void init(const Ctx<TS, NMAX, MIC>& c, St<TS, NMAX, MIC>& st) {
  const int n = c.n, mi = c.mi, m = c.m, l = c.l;
  factor(c, st, true);
  // p = [-q + G^T h; b]: with a2 = 0, a3 = h, sg = 1 the solve forms p1 = v1 + G^T h
  double zero[MIC];
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) zero[cc] = 0.0;
  const double pa = l < n ? -st.qa : (l < m ? -st.qa : 0.0);  // -q  and  b (qa holds -b)
  solve(c, st, pa, zero, st.h, 1.0, st.ds, st.dz);
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

template <typename TS, int NMAX, int MIC>
__device__ __forceinline__ void step(const Ctx<TS, NMAX, MIC>& c, St<TS, NMAX, MIC>& st, bool relax_rule) {
  const double as = ort<MIC>(st.s, st.ds, c.mi, c.l);
  const double az = ort<MIC>(st.z, st.dz, c.mi, c.l);
  const double alpha = relax_rule ? 0.99 * nmin(nmin(1.0, 0.99 * as), 0.99 * az) : 0.99 * nmin(as, az);
  st.xa += alpha * st.da;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    st.s[cc] += alpha * st.ds[cc];
    st.z[cc] += alpha * st.dz[cc];
  }
}

// One predictor-corrector iteration (qpax/pdip.py:87-119); returns the entry convergence flag.
template <typename TS, int NMAX, int MIC>
__device__ __forceinline__ int pc_iter(const Ctx<TS, NMAX, MIC>& c, St<TS, NMAX, MIC>& st, double tol) {
  const int mi = c.mi, l = c.l;
  const double nrm = residuals(c, st, 0.0);
  factor(c, st, false);
  solve(c, st, -st.ra, st.r2, st.r3, -1.0, st.dsa, st.dza);
  // centering (qpax/pdip.py:76-84)
  double sz = 0.0;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) if (64 * cc + l < mi) sz += st.s[cc] * st.z[cc];
  sz = wsum(sz);
  const double mu = sz / mi;
  const double a = nmin(ort<MIC>(st.s, st.dsa, mi, l), ort<MIC>(st.z, st.dza, mi, l));
  double gg = 0.0;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc)
    if (64 * cc + l < mi) gg += (st.s[cc] + a * st.dsa[cc]) * (st.z[cc] + a * st.dza[cc]);
  gg = wsum(gg);
  const double ratio = gg / sz;
  const double sm = ratio * ratio * ratio * mu;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) st.r2[cc] = st.r2[cc] - (sm - st.dsa[cc] * st.dza[cc]);
  solve(c, st, -st.ra, st.r2, st.r3, -1.0, st.ds, st.dz);
  step(c, st, false);
  return nrm < tol ? 1 : 0;
}

// One relaxed Newton iteration (qpax/pdip_relaxed.py:7-52).
template <typename TS, int NMAX, int MIC>
__device__ __forceinline__ int relax_iter(const Ctx<TS, NMAX, MIC>& c, St<TS, NMAX, MIC>& st, double tol,
                                          double kappa) {
  const double nrm = residuals(c, st, kappa);
  factor(c, st, false);
  solve(c, st, -st.ra, st.r2, st.r3, -1.0, st.ds, st.dz);
  step(c, st, true);
  return nrm < tol ? 1 : 0;
}

template <typename TS, int NMAX, int MIC>
__device__ __forceinline__ Ctx<TS, NMAX, MIC> make_ctx(int n, int me, int mi, double* ws) {
  extern __shared__ __align__(16) unsigned char qpx_smem[];
  const Lds w = lds_layout(n, me, mi, (int)sizeof(TS));
  Ctx<TS, NMAX, MIC> c;
  c.n = n; c.me = me; c.mi = mi; c.m = n + me; c.l = threadIdx.x;
  c.Gs = reinterpret_cast<const TS*>(qpx_smem + w.off_G);
  c.As = reinterpret_cast<const TS*>(qpx_smem + w.off_A);
  c.Lp = reinterpret_cast<double*>(qpx_smem + w.off_L);
  c.col = reinterpret_cast<double*>(qpx_smem + w.off_c);
  c.vb = reinterpret_cast<double*>(qpx_smem + w.off_v);
  c.KQ = ws;
  return c;
}

template <typename TS, int NMAX, int MIC>
__device__ __forceinline__ void store_state(const Ctx<TS, NMAX, MIC>& c, const St<TS, NMAX, MIC>& st, int64_t p,
                                            TS* x, TS* s, TS* z, TS* y, bool post) {
  const int n = c.n, me = c.me, mi = c.mi, l = c.l;
  if (l < n) x[p * n + l] = (TS)st.xa;
  else if (l < n + me) y[p * me + (l - n)] = (TS)st.xa;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = 64 * cc + l;
    if (k < mi) {
      double sv = st.s[cc], zv = st.z[cc];
      if (post && st.hinf[cc]) { zv = 0.0; sv = INFINITY; }
      s[p * mi + k] = (TS)sv;
      z[p * mi + k] = (TS)zv;
    }
  }
}

// Forward kernel: solve (and optionally relax) one QP per wavefront.
template <typename TS, int NMAX, int MIC>
__global__ __launch_bounds__(64) void k_wave_solve(Args<TS> a, int with_relax) {
  const int64_t p = blockIdx.x;
  const int n = a.n, me = a.me, mi = a.mi;
  const auto c = make_ctx<TS, NMAX, MIC>(n, me, mi, a.ws + p * a.stride);
  St<TS, NMAX, MIC> st;
  load(c, st, a.Q + p * n * n, a.q + p * n, a.A + p * me * n, a.b + p * me, a.G + p * mi * n, a.h + p * mi);
  init(c, st);
  int conv = 0, it = 0;
  while (it < kMaxIters && conv == 0) {
    conv = pc_iter(c, st, a.tol);
    ++it;
  }
  store_state(c, st, p, a.x, a.s, a.z, a.y, true);
  if (c.l == 0) { a.conv[p] = conv; a.iters[p] = it; }
  if (!with_relax) return;
  // relax from the post-processed point (s = inf, z = 0 on rows with infinite h)
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc)
    if (st.hinf[cc]) { st.z[cc] = 0.0; st.s[cc] = INFINITY; }
  int rconv = 0, rit = 0;
  while (rit < kMaxIters && rconv == 0) {
    rconv = relax_iter(c, st, a.tol, a.kappa);
    ++rit;
  }
  store_state(c, st, p, a.xr, a.sr, a.zr, a.yr, false);
  if (c.l == 0) { a.rconv[p] = rconv; a.riters[p] = rit; }
}

// Backward kernel (qpax/diff_qp.py:20-46 with optnet_derivatives :8-17).
template <typename TS, int NMAX, int MIC>
__global__ __launch_bounds__(64) void k_wave_bwd(Args<TS> a) {
  const int64_t p = blockIdx.x;
  const int n = a.n, me = a.me, mi = a.mi, m = n + me;
  const auto c = make_ctx<TS, NMAX, MIC>(n, me, mi, a.ws + p * a.stride);
  const int l = c.l;
  St<TS, NMAX, MIC> st;
  load(c, st, a.Q + p * n * n, a.q + p * n, a.A + p * me * n, a.b + p * me, a.G + p * mi * n, a.h + p * mi);
  st.xa = l < n ? (double)a.xr_in[p * n + l] : (l < m ? (double)a.yr_in[p * me + (l - n)] : 0.0);
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = 64 * cc + l;
    st.s[cc] = k < mi ? (double)a.sr_in[p * mi + k] : 1.0;
    st.z[cc] = k < mi ? (double)a.zr_in[p * mi + k] : 1.0;
  }
  factor(c, st, false);
  double zero[MIC];
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) zero[cc] = 0.0;
  const double g = l < n ? (double)a.g[p * n + l] : 0.0;
  solve(c, st, -g, zero, zero, 1.0, st.ds, st.dz);
  // cotangents; dz holds dlam_tilde, dlam = dlam_tilde / lam
  c.col[l] = st.xa;
  c.col[64 + l] = st.da;
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = 64 * cc + l;
    if (k < mi) c.vb[k] = st.z[cc];
    if (k < mi) c.Lp[k] = st.dz[cc] / st.z[cc];
  }
  WSYNC();
  const double* X = c.col;
  const double* DX = c.col + 64;
  for (int i = 0; i < n; ++i) {
    for (int j = l; j < n; j += 64)
      a.dQ[(p * n + i) * n + j] = (TS)(0.5 * (DX[i] * X[j] + X[i] * DX[j]));
  }
  if (l < n) a.dq[p * n + l] = (TS)st.da;
  for (int cix = 0; cix < me; ++cix) {
    const double dnu = DX[n + cix], nu = X[n + cix];
    for (int j = l; j < n; j += 64) a.dA[(p * me + cix) * n + j] = (TS)(dnu * X[j] + nu * DX[j]);
  }
  if (l >= n && l < m) a.db[p * me + (l - n)] = (TS)(-st.da);
  for (int k = 0; k < mi; ++k) {
    const double lam = c.vb[k], dlam = c.Lp[k];
    for (int j = l; j < n; j += 64) a.dG[(p * mi + k) * n + j] = (TS)(lam * (dlam * X[j] + lam * DX[j]));
  }
#pragma unroll
  for (int cc = 0; cc < MIC; ++cc) {
    const int k = 64 * cc + l;
    if (k < mi) a.dh[p * mi + k] = (TS)(-st.z[cc] * (st.dz[cc] / st.z[cc]));
  }
}

}  // namespace wave
}  // namespace qpx
