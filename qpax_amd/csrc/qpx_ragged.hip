// Ragged batches: QPs of different sizes (n, m_eq, m_ineq) solved and differentiated in one call.
//
// A plan is built once per size pattern on the host.  It packs the problems' operands back to back
// (element offsets per problem), buckets the problems by the kernel variant that serves their size
// (a wavefront-per-QP instance (NMAX, MIC) or a workgroup-per-QP kernel of NT threads), orders
// each bucket longest-first by its per-iteration flop count, and uploads the size/offset table and
// the orders to the device.  A call then launches one persistent kernel per bucket: its workgroups
// pull problems from the bucket's queue head in the workspace's control block, which a one-block
// fill kernel clears at the start of the call.  A call makes no allocation, no host<->device copy
// and no synchronisation, so it can be captured in a hipGraph.  qpx_plan_tune times the candidate
// launch shapes of every bucket on real data and keeps the fastest.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

#include "../../include/qpx.h"
#include "qpx_args.hpp"
#include "qpx_host.hpp"
#include "qpx_core.hpp"
#include "qpx_wave_launch.hpp"

namespace {

using qpx::RaggedEntry;
using qpx::host::fail;
using qpx::host::kCtlBytes;

__global__ void k_fill_u32(unsigned* p, int count, unsigned v) {
  for (int i = threadIdx.x; i < count; i += blockDim.x) p[i] = v;
}

enum Kind { kWave = 0, kBlock = 1, kGroup = 2 };

struct Bucket {
  int kind;
  qpx::host::WaveVariant wv;  // kWave: kernel instance
  qpx::host::GroupVariant gv; // kGroup: kernel instance
  int nt;                     // kBlock: threads per workgroup
  int wpb;                    // kWave: wavefronts per workgroup
  int lds_per_wave;           // kWave / kGroup: largest LDS image among the bucket's problems
  int persistent;             // 1: grid of resident slots pulling from the queue; 0: one slot per problem
  int64_t first, count;       // range of the device order array
  double flops;               // per-iteration flop estimate of the whole bucket
};

// Per-iteration flops of one predictor-corrector step (residuals, factorisation, two solves).
double iter_flops(int n, int me, int mi) {
  const double N = n, E = me, I = mi;
  const double res = 2 * N * N + 4 * N * I + 4 * N * E;
  const double fac = 2 * I * N * N + N * N * N / 3 + 2 * N * N * E + 2 * N * E * E + E * E * E / 3;
  const double sol = 4 * N * I + 4 * N * N + 4 * N * E + 2 * E * E;
  return res + fac + 2 * sol;
}

int block_threads(int n, int mi) {
  if (n <= 32 && mi <= 128) return 64;
  if (n <= 96) return 128;
  return 256;
}

}  // namespace

struct qpx_plan {
  int dtype = QPX_F32;
  int device = 0;
  int64_t B = 0;
  std::vector<RaggedEntry> table;
  int64_t totals[6] = {0, 0, 0, 0, 0, 0};  // elements of Q, q, A, b, G, h
  int64_t ws_words = 0;
  std::vector<Bucket> buckets;
  std::vector<int> order;  // host copy of the device order
  RaggedEntry* d_table = nullptr;
  int* d_order = nullptr;
  // side streams (default priority, non-blocking) that run buckets beside the caller's stream, and
  // the events that fork them from it and join them back into it; created with the plan
  static constexpr int kSide = 3;
  int n_side = 0;
  hipStream_t side[kSide] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[kSide] = {};
  // held while a call enqueues its fork, bucket launches and joins, so that calls from several
  // host threads on one plan do not interleave their event records and waits
  mutable std::mutex mu;
};

namespace {

size_t plan_ws_bytes(const qpx_plan* p) { return kCtlBytes + (size_t)p->ws_words * sizeof(double); }

template <typename TS>
qpx::Args<TS> bucket_args(const qpx_plan* p, const Bucket& b, int qi, void* ws) {
  qpx::Args<TS> a;
  std::memset(&a, 0, sizeof a);
  a.B = p->B;
  a.counter = reinterpret_cast<unsigned*>(ws);
  a.ws = reinterpret_cast<double*>(static_cast<char*>(ws) + kCtlBytes);
  a.rg = p->d_table;
  a.order = p->d_order + b.first;
  a.order_count = (int)b.count;
  a.qhead = b.persistent ? reinterpret_cast<unsigned*>(ws) + 1 + qi : nullptr;
  a.lds_per_wave = b.lds_per_wave;
  return a;
}

// Workgroups of a launch: the resident slots of the whole device (persistent) or one per problem.
template <typename TS>
unsigned grid_of(const Bucket& b) {
  if (!b.persistent) return (unsigned)(b.kind == kWave ? (b.count + b.wpb - 1) / b.wpb : b.count);
  int per_cu = 0;
  if (b.kind == kWave) {
    per_cu = qpx::wave::occupancy<TS>(b.wv, b.wpb, b.lds_per_wave);
  } else if (b.kind == kGroup) {
    per_cu = qpx::group::occupancy<TS>(b.gv, b.lds_per_wave);
  } else {
    per_cu = b.nt == 64    ? qpx::generic_occupancy<64, TS>(b.lds_per_wave)
             : b.nt == 128 ? qpx::generic_occupancy<128, TS>(b.lds_per_wave)
                           : qpx::generic_occupancy<256, TS>(b.lds_per_wave);
  }
  if (per_cu < 1) per_cu = 1;
  const int64_t slots = (int64_t)per_cu * qpx::host::cu_count();
  const int64_t per_slot = b.kind == kWave ? b.wpb : 1;
  const int64_t need = (b.count + per_slot - 1) / per_slot;
  return (unsigned)std::min(slots, need);
}

template <typename TS>
hipError_t launch_bucket_fwd(const qpx::Args<TS>& a, const Bucket& b, hipStream_t st, int relax) {
  const unsigned grid = grid_of<TS>(b);
  if (b.kind == kWave) return qpx::wave::launch_solve(b.wv, a, grid, b.wpb, st, relax);
  if (b.kind == kGroup) return qpx::group::launch_solve(b.gv, a, grid, b.lds_per_wave, st, relax);
  if (b.nt == 64) return qpx::launch_generic_solve<64, TS>(a, grid, b.lds_per_wave, st, relax);
  if (b.nt == 128) return qpx::launch_generic_solve<128, TS>(a, grid, b.lds_per_wave, st, relax);
  return qpx::launch_generic_solve<256, TS>(a, grid, b.lds_per_wave, st, relax);
}

template <typename TS>
hipError_t launch_bucket_bwd(const qpx::Args<TS>& a, const Bucket& b, hipStream_t st) {
  const unsigned grid = grid_of<TS>(b);
  if (b.kind == kWave) return qpx::wave::launch_bwd(b.wv, a, grid, b.wpb, st);
  if (b.kind == kGroup) return qpx::group::launch_bwd(b.gv, a, grid, b.lds_per_wave, st);
  if (b.nt == 64) return qpx::launch_generic_bwd<64, TS>(a, grid, b.lds_per_wave, st);
  if (b.nt == 128) return qpx::launch_generic_bwd<128, TS>(a, grid, b.lds_per_wave, st);
  return qpx::launch_generic_bwd<256, TS>(a, grid, b.lds_per_wave, st);
}

struct FwdIO {
  const void *Q, *q, *A, *b, *G, *h;
  double tol, kappa;
  void *x, *s, *z, *y;
  int32_t *conv, *iters;
  void *xr, *sr, *zr, *yr;
  int32_t *rconv, *riters;
};

struct BwdIO {
  const void *Q, *q, *A, *b, *G, *h, *xr, *sr, *zr, *yr, *g;
  int32_t flags;
  void *dQ, *dq, *dA, *db, *dG, *dh;
};

template <typename TS>
void fill_fwd(qpx::Args<TS>& a, const FwdIO& io, bool relax) {
  a.Q = (const TS*)io.Q; a.q = (const TS*)io.q; a.A = (const TS*)io.A; a.b = (const TS*)io.b;
  a.G = (const TS*)io.G; a.h = (const TS*)io.h;
  a.x = (TS*)io.x; a.s = (TS*)io.s; a.z = (TS*)io.z; a.y = (TS*)io.y; a.conv = io.conv; a.iters = io.iters;
  if (relax) {
    a.xr = (TS*)io.xr; a.sr = (TS*)io.sr; a.zr = (TS*)io.zr; a.yr = (TS*)io.yr;
    a.rconv = io.rconv; a.riters = io.riters;
  }
  a.tol = io.tol; a.kappa = io.kappa;
}

template <typename TS>
void fill_bwd(qpx::Args<TS>& a, const BwdIO& io) {
  a.Q = (const TS*)io.Q; a.q = (const TS*)io.q; a.A = (const TS*)io.A; a.b = (const TS*)io.b;
  a.G = (const TS*)io.G; a.h = (const TS*)io.h;
  a.xr_in = (const TS*)io.xr; a.sr_in = (const TS*)io.sr; a.zr_in = (const TS*)io.zr; a.yr_in = (const TS*)io.yr;
  a.g = (const TS*)io.g; a.flags = io.flags;
  a.dQ = (TS*)io.dQ; a.dq = (TS*)io.dq; a.dA = (TS*)io.dA; a.db = (TS*)io.db; a.dG = (TS*)io.dG; a.dh = (TS*)io.dh;
}

hipError_t reset_queues(const qpx_plan* p, void* ws, hipStream_t st) {
  const int nq = (int)p->buckets.size();
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, st, reinterpret_cast<unsigned*>(ws) + 1, nq, 0u);
  return hipGetLastError();
}

// Launches every bucket, bucket i on lane i % lanes: lane 0 is the caller's stream, the others the
// plan's side streams, forked from the caller's stream after the queue reset and joined back into it
// at the end.  Buckets share no memory (own queue head, own problems' outputs and workspace slices),
// so their order does not change any result.  Run side by side, a persistent workgroup that finds
// its bucket's queue empty frees its CU for the workgroups of a bucket on another lane, instead of
// each bucket's tail leaving CUs idle before the next launch starts.  Event records and waits are
// legal under stream capture: a captured call becomes a graph with one branch per lane.
template <typename L>
qpx_status fan_out(const qpx_plan* p, hipStream_t st, L&& launch) {
  const int nb = (int)p->buckets.size();
  const int lanes = std::min(nb, 1 + p->n_side);
  if (lanes <= 1) {
    for (int i = 0; i < nb; ++i) QPX_HIP(launch(i, st));
    return QPX_OK;
  }
  QPX_HIP(hipEventRecord(p->fork, st));
  for (int k = 1; k < lanes; ++k) QPX_HIP(hipStreamWaitEvent(p->side[k - 1], p->fork, 0));
  qpx_status st_launch = QPX_OK;
  for (int i = 0; i < nb && st_launch == QPX_OK; ++i) {
    const int k = i % lanes;
    if (launch(i, k == 0 ? st : p->side[k - 1]) != hipSuccess)
      st_launch = fail(QPX_ERR_HIP, "launch of bucket %d failed: %s", i, hipGetErrorString(hipGetLastError()));
  }
  // join every forked lane, also after a failed launch, so that a capture is never left open on a
  // side stream
  for (int k = 1; k < lanes; ++k) {
    QPX_HIP(hipEventRecord(p->join[k - 1], p->side[k - 1]));
    QPX_HIP(hipStreamWaitEvent(st, p->join[k - 1], 0));
  }
  return st_launch;
}

template <typename TS>
qpx_status run_fwd(const qpx_plan* p, const FwdIO& io, bool relax, void* ws, hipStream_t st, int only = -1) {
  std::lock_guard<std::mutex> lock(p->mu);
  QPX_HIP(reset_queues(p, ws, st));
  auto launch = [&](int i, hipStream_t s) {
    const Bucket& b = p->buckets[i];
    auto a = bucket_args<TS>(p, b, i, ws);
    fill_fwd(a, io, relax);
    return launch_bucket_fwd<TS>(a, b, s, relax ? 1 : 0);
  };
  if (only >= 0) {
    QPX_HIP(launch(only, st));
    return QPX_OK;
  }
  return fan_out(p, st, launch);
}

template <typename TS>
qpx_status run_bwd(const qpx_plan* p, const BwdIO& io, void* ws, hipStream_t st) {
  std::lock_guard<std::mutex> lock(p->mu);
  QPX_HIP(reset_queues(p, ws, st));
  return fan_out(p, st, [&](int i, hipStream_t s) {
    const Bucket& b = p->buckets[i];
    auto a = bucket_args<TS>(p, b, i, ws);
    fill_bwd(a, io);
    return launch_bucket_bwd<TS>(a, b, s);
  });
}

qpx_status check_call(const qpx_plan* p, size_t ws_bytes, void* ws) {
  if (!p) return fail(QPX_ERR_ARG, "null plan");
  if (ws_bytes < plan_ws_bytes(p))
    return fail(QPX_ERR_WORKSPACE, "workspace of %zu bytes is smaller than the %zu required", ws_bytes, plan_ws_bytes(p));
  if (p->B > 0 && !ws) return fail(QPX_ERR_ARG, "null workspace");
  int dev = -1;
  QPX_HIP(hipGetDevice(&dev));
  if (dev != p->device) return fail(QPX_ERR_ARG, "plan was made on device %d, current device is %d", p->device, dev);
  return QPX_OK;
}

bool any_null(std::initializer_list<const void*> ps) {
  for (auto q : ps)
    if (!q) return true;
  return false;
}

qpx_status check_fwd(const qpx_plan* p, const FwdIO& io, bool relax) {
  if (p->B == 0) return QPX_OK;
  if (any_null({io.Q, io.q, io.G, io.h, io.x, io.s, io.z, io.conv, io.iters}))
    return fail(QPX_ERR_ARG, "null pointer argument");
  if (p->totals[3] > 0 && any_null({io.A, io.b, io.y})) return fail(QPX_ERR_ARG, "null equality pointer");
  if (relax && any_null({io.xr, io.sr, io.zr, io.rconv, io.riters})) return fail(QPX_ERR_ARG, "null relaxed output");
  if (relax && p->totals[3] > 0 && !io.yr) return fail(QPX_ERR_ARG, "null equality pointer");
  return QPX_OK;
}

template <typename F>
qpx_status by_dtype(const qpx_plan* p, F&& f) {
  if (p->dtype == QPX_F32) return f(float());
  return f(double());
}

}  // namespace

extern "C" {

QPX_API qpx_status qpx_plan_create(int32_t dtype, int64_t batch, const int32_t* n, const int32_t* m_eq,
                                   const int32_t* m_ineq, qpx_plan** out) {
  qpx::host::clear_error();
  if (!out) return fail(QPX_ERR_ARG, "null plan output");
  *out = nullptr;
  if (dtype != QPX_F32 && dtype != QPX_F64) return fail(QPX_ERR_ARG, "dtype must be QPX_F32 or QPX_F64 (got %d)", dtype);
  if (batch < 0 || batch > 0x7fffffffLL) return fail(QPX_ERR_ARG, "invalid batch %lld", (long long)batch);
  if (batch > 0 && (!n || !m_eq || !m_ineq)) return fail(QPX_ERR_ARG, "null size array");
  const int ts = dtype == QPX_F32 ? 4 : 8;
  auto* p = new qpx_plan;
  p->dtype = dtype;
  p->B = batch;
  if (hipGetDevice(&p->device) != hipSuccess) {
    delete p;
    return fail(QPX_ERR_HIP, "hipGetDevice failed");
  }
  p->table.resize(batch);
  int64_t o[6] = {0, 0, 0, 0, 0, 0}, ows = 0;
  for (int64_t i = 0; i < batch; ++i) {
    const int ni = n[i], ei = m_eq[i], ii = m_ineq[i];
    if (ni < 1 || ei < 0 || ii < 1) {
      delete p;
      return fail(QPX_ERR_ARG, "problem %lld: invalid sizes n=%d m_eq=%d m_ineq=%d (n >= 1, m_ineq >= 1)",
                  (long long)i, ni, ei, ii);
    }
    RaggedEntry& e = p->table[i];
    e.n = ni; e.me = ei; e.mi = ii; e.reserved = 0;
    e.off_Q = o[0]; e.off_q = o[1]; e.off_A = o[2]; e.off_b = o[3]; e.off_G = o[4]; e.off_h = o[5]; e.off_ws = ows;
    o[0] += (int64_t)ni * ni; o[1] += ni; o[2] += (int64_t)ei * ni; o[3] += ei; o[4] += (int64_t)ii * ni; o[5] += ii;
    ows += qpx::host::stride_words(ni, ei, ii);
  }
  std::copy(o, o + 6, p->totals);
  p->ws_words = ows;

  // buckets: key = kernel variant
  std::vector<std::vector<int>> members;
  std::vector<Bucket> bks;
  auto key_of = [&](const RaggedEntry& e, Bucket* b) {
    std::memset(b, 0, sizeof *b);
    qpx::host::WaveVariant wv;
    qpx::host::GroupVariant gv;
    if (qpx::host::wave_variant(e.n, e.me, e.mi, ts, &wv)) {
      b->kind = kWave; b->wv = wv; b->wpb = 1;
    } else if (qpx::host::group_variant(e.n, e.me, e.mi, ts, 0, &gv)) {
      b->kind = kGroup; b->gv = gv;
    } else {
      b->kind = kBlock; b->nt = block_threads(e.n, e.mi);
    }
  };
  for (int64_t i = 0; i < batch; ++i) {
    Bucket k;
    key_of(p->table[i], &k);
    size_t j = 0;
    for (; j < bks.size(); ++j)
      if (bks[j].kind == k.kind && bks[j].nt == k.nt && bks[j].wv.nmax == k.wv.nmax && bks[j].wv.mic == k.wv.mic &&
          bks[j].gv.nti == k.gv.nti && bks[j].gv.w == k.gv.w)
        break;
    if (j == bks.size()) { bks.push_back(k); members.emplace_back(); }
    members[j].push_back((int)i);
  }
  if (bks.size() > (size_t)qpx::host::kMaxQueues) {
    delete p;
    return fail(QPX_ERR_UNSUPPORTED, "%zu kernel buckets exceed the %d queue heads", bks.size(), qpx::host::kMaxQueues);
  }
  std::vector<int> order;
  order.reserve(batch);
  for (size_t j = 0; j < bks.size(); ++j) {
    auto& mem = members[j];
    std::vector<double> cost(mem.size());
    double total = 0.0;
    int lds = 0;
    for (size_t k = 0; k < mem.size(); ++k) {
      const RaggedEntry& e = p->table[mem[k]];
      cost[k] = iter_flops(e.n, e.me, e.mi);
      total += cost[k];
      if (bks[j].kind == kWave) lds = std::max(lds, qpx::wave::lds_layout(e.n, e.me, e.mi, ts).bytes);
      if (bks[j].kind == kGroup) lds = std::max(lds, qpx::host::group_lds_bytes(e.n, e.me, e.mi, ts, bks[j].gv));
      if (bks[j].kind == kBlock) lds = std::max(lds, qpx::generic_lds_bytes(e.n, e.me, bks[j].nt));
    }
    std::vector<size_t> idx(mem.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
    bks[j].first = (int64_t)order.size();
    bks[j].count = (int64_t)mem.size();
    bks[j].flops = total;
    bks[j].lds_per_wave = lds;
    bks[j].persistent = 1;
    for (size_t k : idx) order.push_back(mem[k]);
  }
  // most expensive bucket first
  std::vector<size_t> bo(bks.size());
  std::iota(bo.begin(), bo.end(), 0);
  std::stable_sort(bo.begin(), bo.end(), [&](size_t a, size_t b) { return bks[a].flops > bks[b].flops; });
  for (size_t k : bo) p->buckets.push_back(bks[k]);
  p->order = order;

  if (batch > 0) {
    if (hipMalloc(reinterpret_cast<void**>(&p->d_table), sizeof(RaggedEntry) * batch) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&p->d_order), sizeof(int) * batch) != hipSuccess ||
        hipMemcpy(p->d_table, p->table.data(), sizeof(RaggedEntry) * batch, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->d_order, order.data(), sizeof(int) * batch, hipMemcpyHostToDevice) != hipSuccess) {
      if (p->d_table) (void)hipFree(p->d_table);
      if (p->d_order) (void)hipFree(p->d_order);
      delete p;
      return fail(QPX_ERR_HIP, "allocating or uploading the plan's device tables failed");
    }
  }
  // lanes for the buckets: QPX_RAGGED_LANES (1 = every bucket on the caller's stream), default
  // 1 + kSide, never more than there are buckets
  int lanes = 1 + qpx_plan::kSide;
  if (const char* env = std::getenv("QPX_RAGGED_LANES")) lanes = std::atoi(env);
  lanes = std::max(1, std::min({lanes, 1 + qpx_plan::kSide, (int)p->buckets.size()}));
  bool ok = hipEventCreateWithFlags(&p->fork, hipEventDisableTiming) == hipSuccess;
  for (int k = 0; ok && k < lanes - 1; ++k) {
    ok = hipStreamCreateWithFlags(&p->side[k], hipStreamNonBlocking) == hipSuccess;
    if (ok && hipEventCreateWithFlags(&p->join[k], hipEventDisableTiming) != hipSuccess) {
      (void)hipStreamDestroy(p->side[k]);
      ok = false;
    }
    if (ok) p->n_side = k + 1;
  }
  if (!ok) {
    qpx_plan_destroy(p);
    return fail(QPX_ERR_HIP, "creating the plan's side streams or events failed");
  }
  *out = p;
  return QPX_OK;
}

QPX_API void qpx_plan_destroy(qpx_plan* p) {
  if (!p) return;
  int cur = -1;
  (void)hipGetDevice(&cur);
  if (cur != p->device) (void)hipSetDevice(p->device);
  if (p->d_table) (void)hipFree(p->d_table);
  if (p->d_order) (void)hipFree(p->d_order);
  for (int k = 0; k < p->n_side; ++k) {
    (void)hipEventDestroy(p->join[k]);
    (void)hipStreamDestroy(p->side[k]);
  }
  if (p->fork) (void)hipEventDestroy(p->fork);
  if (cur >= 0 && cur != p->device) (void)hipSetDevice(cur);
  delete p;
}

QPX_API qpx_status qpx_plan_info(const qpx_plan* p, int64_t* info) {
  qpx::host::clear_error();
  if (!p || !info) return fail(QPX_ERR_ARG, "null argument");
  for (int k = 0; k < 6; ++k) info[k] = p->totals[k];
  info[6] = (int64_t)plan_ws_bytes(p);
  info[7] = (int64_t)p->buckets.size();
  return QPX_OK;
}

QPX_API qpx_status qpx_plan_offsets(const qpx_plan* p, int64_t* off) {
  qpx::host::clear_error();
  if (!p || (!off && p->B > 0)) return fail(QPX_ERR_ARG, "null argument");
  for (int64_t i = 0; i < p->B; ++i) {
    const RaggedEntry& e = p->table[i];
    int64_t* r = off + 6 * i;
    r[0] = e.off_Q; r[1] = e.off_q; r[2] = e.off_A; r[3] = e.off_b; r[4] = e.off_G; r[5] = e.off_h;
  }
  return QPX_OK;
}

QPX_API qpx_status qpx_plan_bucket(const qpx_plan* p, int32_t i, int64_t* out) {
  qpx::host::clear_error();
  if (!p || !out) return fail(QPX_ERR_ARG, "null argument");
  if (i < 0 || i >= (int)p->buckets.size()) return fail(QPX_ERR_ARG, "bucket %d out of range", i);
  const Bucket& b = p->buckets[i];
  out[0] = b.kind; out[1] = b.kind == kGroup ? b.gv.nti : b.wv.nmax; out[2] = b.wv.mic;
  out[3] = b.kind == kGroup ? 64 * b.gv.w : b.nt; out[4] = b.kind == kGroup ? b.gv.w : b.wpb; out[5] = b.persistent;
  out[6] = b.count; out[7] = b.lds_per_wave;
  out[8] = p->dtype == QPX_F32 ? grid_of<float>(b) : grid_of<double>(b);
  out[9] = (int64_t)b.flops;
  return QPX_OK;
}

QPX_API qpx_status qpx_plan_solve(const qpx_plan* p, const void* Q, const void* q, const void* A, const void* b,
                                  const void* G, const void* h, double solver_tol, void* x, void* s, void* z, void* y,
                                  int32_t* converged, int32_t* iters, void* workspace, size_t workspace_bytes,
                                  void* stream) {
  qpx::host::clear_error();
  qpx_status st = check_call(p, workspace_bytes, workspace);
  if (st != QPX_OK) return st;
  FwdIO io{Q, q, A, b, G, h, solver_tol, 0.0, x, s, z, y, converged, iters,
           nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  if ((st = check_fwd(p, io, false)) != QPX_OK || p->B == 0) return st;
  return by_dtype(p, [&](auto t) { return run_fwd<decltype(t)>(p, io, false, workspace, (hipStream_t)stream); });
}

QPX_API qpx_status qpx_plan_solve_relaxed(const qpx_plan* p, const void* Q, const void* q, const void* A,
                                          const void* b, const void* G, const void* h, double solver_tol,
                                          double target_kappa, void* x, void* s, void* z, void* y,
                                          int32_t* converged, int32_t* iters, void* xr, void* sr, void* zr,
                                          void* yr, int32_t* relax_converged, int32_t* relax_iters,
                                          void* workspace, size_t workspace_bytes, void* stream) {
  qpx::host::clear_error();
  qpx_status st = check_call(p, workspace_bytes, workspace);
  if (st != QPX_OK) return st;
  FwdIO io{Q, q, A, b, G, h, solver_tol, target_kappa, x, s, z, y, converged, iters,
           xr, sr, zr, yr, relax_converged, relax_iters};
  if ((st = check_fwd(p, io, true)) != QPX_OK || p->B == 0) return st;
  return by_dtype(p, [&](auto t) { return run_fwd<decltype(t)>(p, io, true, workspace, (hipStream_t)stream); });
}

QPX_API qpx_status qpx_plan_diff(const qpx_plan* p, const void* Q, const void* q, const void* A, const void* b,
                                 const void* G, const void* h, const void* xr, const void* sr, const void* zr,
                                 const void* yr, const void* dl_dx, int32_t flags, void* dQ, void* dq, void* dA,
                                 void* db, void* dG, void* dh, void* workspace, size_t workspace_bytes,
                                 void* stream) {
  qpx::host::clear_error();
  qpx_status st = check_call(p, workspace_bytes, workspace);
  if (st != QPX_OK || p->B == 0) return st;
  if (any_null({Q, q, G, h, xr, sr, zr, dl_dx, dQ, dq, dG, dh})) return fail(QPX_ERR_ARG, "null pointer argument");
  if (p->totals[3] > 0 && any_null({A, b, yr, dA, db})) return fail(QPX_ERR_ARG, "null equality pointer");
  BwdIO io{Q, q, A, b, G, h, xr, sr, zr, yr, dl_dx, flags, dQ, dq, dA, db, dG, dh};
  return by_dtype(p, [&](auto t) { return run_bwd<decltype(t)>(p, io, workspace, (hipStream_t)stream); });
}

// Times every candidate launch shape of every bucket on the given problems (a full solve + relax
// of that bucket per candidate, after one warm-up run) and keeps the fastest.  Synchronises the
// stream; the outputs hold the last candidate's (identical) results afterwards.
QPX_API qpx_status qpx_plan_tune(qpx_plan* p, const void* Q, const void* q, const void* A, const void* b,
                                 const void* G, const void* h, double solver_tol, double target_kappa, void* x,
                                 void* s, void* z, void* y, int32_t* converged, int32_t* iters, void* xr, void* sr,
                                 void* zr, void* yr, int32_t* relax_converged, int32_t* relax_iters,
                                 void* workspace, size_t workspace_bytes, void* stream) {
  qpx::host::clear_error();
  qpx_status st = check_call(p, workspace_bytes, workspace);
  if (st != QPX_OK) return st;
  FwdIO io{Q, q, A, b, G, h, solver_tol, target_kappa, x, s, z, y, converged, iters,
           xr, sr, zr, yr, relax_converged, relax_iters};
  if ((st = check_fwd(p, io, true)) != QPX_OK || p->B == 0) return st;
  hipStream_t hs = (hipStream_t)stream;
  const int ts = p->dtype == QPX_F32 ? 4 : 8;
  const std::vector<int>& host_order = p->order;
  hipEvent_t e0, e1;
  QPX_HIP(hipEventCreate(&e0));
  QPX_HIP(hipEventCreate(&e1));
  for (int i = 0; i < (int)p->buckets.size(); ++i) {
    Bucket& bk = p->buckets[i];
    std::vector<Bucket> cands;
    if (bk.kind == kWave) {
      for (int wpb : {1, 2, 4})
        for (int pers : {1, 0}) {
          if ((int64_t)wpb * bk.lds_per_wave > qpx::host::max_lds_per_block()) continue;
          Bucket c = bk;
          c.wpb = wpb;
          c.persistent = pers;
          cands.push_back(c);
        }
    } else if (bk.kind == kGroup) {
      for (int w : {4, 8})
        for (int pers : {1, 0}) {
          Bucket c = bk;
          c.gv.w = w;
          int lds = 0;
          for (int64_t k = 0; k < bk.count; ++k) {
            const RaggedEntry& e = p->table[host_order[bk.first + k]];
            lds = std::max(lds, qpx::host::group_lds_bytes(e.n, e.me, e.mi, ts, c.gv));
          }
          if (lds + 256 > qpx::host::max_lds_per_block()) continue;
          c.lds_per_wave = lds;
          c.persistent = pers;
          cands.push_back(c);
        }
    } else {
      for (int nt : {64, 128, 256})
        for (int pers : {1, 0}) {
          Bucket c = bk;
          c.nt = nt;
          c.persistent = pers;
          cands.push_back(c);
        }
    }
    (void)ts;
    float best = 1e30f;
    Bucket keep = bk;
    for (const Bucket& c : cands) {
      bk = c;
      float ms = 0.f;
      for (int rep = 0; rep < 2; ++rep) {
        st = by_dtype(p, [&](auto t) { return run_fwd<decltype(t)>(p, io, true, workspace, hs, i); });
        if (st != QPX_OK) break;
        if (rep == 0) {
          QPX_HIP(hipEventRecord(e0, hs));
        } else {
          QPX_HIP(hipEventRecord(e1, hs));
          QPX_HIP(hipEventSynchronize(e1));
          QPX_HIP(hipEventElapsedTime(&ms, e0, e1));
        }
      }
      if (st != QPX_OK) break;
      if (ms < best) { best = ms; keep = c; }
    }
    bk = keep;
    if (st != QPX_OK) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st != QPX_OK) return st;
  return by_dtype(p, [&](auto t) { return run_fwd<decltype(t)>(p, io, true, workspace, hs); });
}

QPX_API size_t qpx_plan_workspace_size(const qpx_plan* p) { return p ? plan_ws_bytes(p) : 0; }

}  // extern "C"
