"""Benchmark: batched QP forward+backward throughput on one MI355X.

Workload (the headline metric's configuration): 4096 random feasible dense QPs, n=50, m_eq=10,
m_ineq=100, float32 inputs/outputs; one step = qpax_amd.solve_qp_primal forward (solve + kappa
relaxation) and backward of loss = ||x - 1||^2 at solver_tol=1e-5, target_kappa=1e-3, i.e. the
gradients w.r.t. Q, q, A, b, G, h.  The fused path (one launch for solve+relax, one for the
backward) is timed for --steps steps; the plain per-stage path is timed on --plain-steps steps for
the fused/plain ratio.  Inputs are resident in HBM before timing starts.

Prints ONE JSON line (rank 0).  Extra objects: "roofline" for the dominant kernel (fused solve +
relax) from HIP events on the launch stream, and "cpu_baseline" (the float64 numpy oracle on the
same workload, one single-threaded process per host core, with the CPU model and core counts).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

N, ME, MI, BATCH, SEED = 50, 10, 100, 4096, 3
TOL, KAPPA = 1e-5, 1e-3
PEAK_F64_TFLOPS = 78.6  # MI355X float64 (vector and matrix) dense peak
PEAK_HBM_GBS = 8000.0


def flops_per_qp(n, me, mi, it_fwd, it_rx, bwd=True):
    """Algorithmic flops of one QP (multiply-add = 2; dense, non-symmetric G^T D G; Cholesky)."""
    f_res = 2 * n * n + 4 * n * mi + 4 * n * me
    f_fac = 2 * mi * n * n + n ** 3 / 3 + 2 * n * n * me + 2 * n * me * me + me ** 3 / 3
    f_sol = 4 * n * mi + 4 * n * n + 4 * n * me + 2 * me * me
    f_pc = f_res + f_fac + 2 * f_sol
    f_rx = f_res + f_fac + f_sol
    f_init = f_fac + f_sol + 2 * n * mi
    f_bwd = f_fac + f_sol + 3 * n * n + 2 * n * me + 3 * n * mi
    fwd = f_init + it_fwd * f_pc + it_rx * f_rx
    return fwd, (f_bwd if bwd else 0.0)


def bytes_per_qp(n, me, mi, isz=4):
    inputs = n * n + n + me * n + me + mi * n + mi
    fwd_out = 2 * (n + 2 * mi + me) * isz + 16  # x,s,z,y and xr,sr,zr,yr + flags
    return inputs * isz, fwd_out, (n + inputs) * isz


def profiled_traffic(profile, key):
    """HBM bytes per launch recorded as `traffic_bytes_per_launch <key>: <bytes>` in the committed
    rocprofv3 summary profiles/<profile> (FETCH_SIZE x2 + WRITE_SIZE, calibrated on a plain
    copy), or None."""
    path = os.path.join(HERE, "profiles", profile)
    try:
        for line in open(path):
            if line.startswith(f"traffic_bytes_per_launch {key}:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def hbm_fields(traffic, ms):
    """Achieved HBM bandwidth of `traffic` bytes moved in `ms` milliseconds, and its fraction of peak."""
    if traffic is None or ms <= 0:
        return {"hbm_GBs": None, "hbm_frac": None}
    gbs = traffic / (ms * 1e-3) / 1e9
    return {"hbm_GBs": round(gbs, 1), "hbm_frac": round(gbs / PEAK_HBM_GBS, 4)}


def make_inputs(torch, dev):
    from tests import problems as P

    data = [np.asarray(v, np.float32) for v in P.random_feasible(BATCH, N, ME, MI, seed=SEED)]
    return data, [torch.tensor(v, device=dev, requires_grad=True) for v in data]


def fwd_bwd(qpax_amd, ts, mode):
    for t in ts:
        t.grad = None
    x = qpax_amd.solve_qp_primal(*ts, solver_tol=TOL, target_kappa=KAPPA, mode=mode)
    loss = ((x - 1.0) ** 2).sum()
    loss.backward()
    return x


def time_steps(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / max(steps, 1)


def kernel_times(torch, qpax_amd, ts, reps):
    """Average duration of the fused solve+relax launch and of the fused backward launch, from HIP
    events recorded on the stream the library launches on (torch's current stream)."""
    from qpax_amd._tensors import normalize
    from qpax_amd.diff_qp import diff_qp_batch
    from qpax_amd.pdip import solve_qp_relaxed_batch

    bt = normalize(*[t.detach() for t in ts])
    st = torch.cuda.current_stream()
    f_ms, b_ms = [], []
    out = None
    for r in range(reps + 1):
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        e0.record(st)
        out = solve_qp_relaxed_batch(bt, TOL, KAPPA, "fused", check=False)
        e1.record(st)
        g = 2.0 * (out[0] - 1.0)
        e1b = torch.cuda.Event(enable_timing=True)
        e1b.record(st)
        diff_qp_batch(bt, out[6:10], g, "fused")
        e2.record(st)
        torch.cuda.synchronize()
        if r > 0:
            f_ms.append(e0.elapsed_time(e1))
            b_ms.append(e1b.elapsed_time(e2))
    return float(np.mean(f_ms)), float(np.mean(b_ms)), out


_CPU_DATA = None  # the batch, inherited by the forked workers of cpu_baseline


def _cpu_chunk(rng):
    """Oracle fwd (solve + relax) + bwd on problems rng[0]:rng[1] of _CPU_DATA (one worker)."""
    from threadpoolctl import threadpool_limits

    from oracle import qp_oracle as O

    lo, hi = rng
    d64 = [v[lo:hi].astype(np.float64) for v in _CPU_DATA]
    with threadpool_limits(1):
        xo, rel, info = O.solve_qp_primal_fwd(*d64, solver_tol=TOL, target_kappa=KAPPA)
        go = O.solve_qp_primal_bwd(*d64, rel, 2 * (xo - 1))
    return xo, go, info


def host_cpu():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return model, os.cpu_count() or 1, affinity


def cpu_baseline(data, sample):
    """float64 numpy oracle on the first `sample` problems, fwd (solve + relax) + bwd, split over
    one forked worker per host core of this process's affinity (at most 16, the GPU's share of the
    host).  Runs before anything touches the GPU, so the workers are plain forks."""
    import multiprocessing as mp

    global _CPU_DATA
    model, nproc, affinity = host_cpu()
    workers = max(1, min(16, affinity, sample))
    _CPU_DATA = data
    bounds = np.linspace(0, sample, workers + 1).astype(int)
    chunks = [(int(a), int(b)) for a, b in zip(bounds[:-1], bounds[1:]) if b > a]
    with mp.get_context("fork").Pool(len(chunks)) as pool:
        pool.map(_cpu_chunk, [(a, min(a + 1, b)) for a, b in chunks])  # warm-up (imports)
        t0 = time.perf_counter()
        parts = pool.map(_cpu_chunk, chunks)
        dt = time.perf_counter() - t0
    _CPU_DATA = None
    xo = np.concatenate([p[0] for p in parts])
    go = [np.concatenate([p[1][k] for p in parts]) for k in range(6)]
    info = [np.concatenate([p[2][k] for p in parts]) for k in range(4)]
    host = {"cpu_model": model, "nproc": nproc, "affinity_cores": affinity, "workers": len(chunks)}
    return sample / dt, xo, go, info, host


def uneven_bench(torch, args):
    """Uneven batching: 16384 QPs with n ~ U{8..128}, m_ineq ~ U{n..3n}, m_eq ~ U{0..n/4} (seed 4),
    float32 fwd (solve + relax) + bwd through one autotuned ragged plan, captured in a CUDA graph
    and replayed; one step = one replay."""
    from tests import problems as P
    from qpax_amd.ragged import RaggedPlan

    B = args.uneven_batch
    probs = P.uneven(B, seed=4)
    n = np.array([p[1].shape[0] for p in probs])
    me = np.array([p[3].shape[0] for p in probs])
    mi = np.array([p[5].shape[0] for p in probs])
    plan = RaggedPlan(n, me, mi, dtype=torch.float32)
    ops = plan.pack(probs)
    t0 = time.perf_counter()
    plan.tune(ops, TOL, KAPPA)
    torch.cuda.synchronize()
    tune_s = time.perf_counter() - t0
    out = plan.solve_relaxed(ops, TOL, KAPPA)
    dl = (2.0 * (out[0] - 1.0)).contiguous()
    g, fwd, grads = plan.capture(ops, dl, TOL, KAPPA)
    ms = 1e3 * time_steps(torch, g.replay, args.steps, args.warmup)
    # forward alone (eager) for the roofline of the solve + relax launches
    st = torch.cuda.current_stream()
    f_ms = []
    for r in range(4):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        plan.solve_relaxed(ops, TOL, KAPPA, out=fwd)
        e1.record(st)
        torch.cuda.synchronize()
        if r:
            f_ms.append(e0.elapsed_time(e1))
    f_ms = float(np.mean(f_ms))
    it_f = fwd[5].cpu().numpy().astype(np.float64)
    it_r = fwd[11].cpu().numpy().astype(np.float64)
    fl = sum(flops_per_qp(a, b, c, d, e, bwd=False)[0] for a, b, c, d, e in zip(n, me, mi, it_f, it_r))
    achieved = fl / (f_ms * 1e-3) / 1e12
    res = {
        "metric": "fwd+bwd QPs/s, uneven batch (hipGraph-captured ragged plan)",
        "value": round(B / (ms * 1e-3), 1), "unit": "QPs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{B} QPs, n~U{{8..128}}, m_ineq~U{{n..3n}}, m_eq~U{{0..n/4}}, seed 4, float32 "
                               "inputs/outputs, float64 arithmetic, solve_qp_primal fwd+bwd, tol 1e-5, kappa 1e-3",
                   "batch": B, "io_dtype": "f32"},
        "tune_s": round(tune_s, 2),
        "fwd_ms": round(f_ms, 3),
        "converged_frac": float((fwd[4] == 1).float().mean()),
        "relax_converged_frac": float((fwd[10] == 1).float().mean()),
        "buckets": plan.buckets(),
        "roofline": {"kernel": "all buckets of the forward (solve + relax)", "bound": "mfma",
                     "achieved": round(achieved, 3), "peak": PEAK_F64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_F64_TFLOPS, 4),
                     "traffic": profiled_traffic("uneven_16384_fp32.md", "forward"),
                     "traffic_source": "profiles/uneven_16384_fp32.md (rocprofv3 FETCH_SIZE x2 + WRITE_SIZE, "
                                       "summed over the forward's bucket launches)",
                     **hbm_fields(profiled_traffic("uneven_16384_fp32.md", "forward"), f_ms),
                     "algorithmic_flops_per_launch": fl},
    }
    if not args.no_cpu:
        from oracle import qp_oracle as O

        rng = np.random.default_rng(0)
        idx = rng.choice(B, args.uneven_check, replace=False)
        ex, eg = [], []
        x = fwd[0]
        for i in idx:
            d = [np.asarray(np.asarray(v, np.float32), np.float64)[None] for v in probs[i]]
            xo, rel, info = O.solve_qp_primal_fwd(*d, solver_tol=TOL, target_kappa=KAPPA)
            if info[0][0] != 1 or info[2][0] != 1:
                continue
            go = O.solve_qp_primal_bwd(*d, rel, 2 * (xo - 1))
            xg = plan.unpack(x, "q", int(i)).cpu().numpy().astype(np.float64)
            ex.append(np.linalg.norm(xg - xo[0]) / max(np.linalg.norm(xo[0]), 1e-12))
            for k, w in enumerate("Q q A b G h".split()):
                if go[k].size:
                    gg = plan.unpack(grads[k], w, int(i)).cpu().numpy().astype(np.float64)
                    eg.append(np.linalg.norm(gg - go[k][0]) / max(np.linalg.norm(go[k][0]), 1e-12))
        res["x_rel_err_vs_f64_oracle"] = {"max": float(max(ex)), "median": float(np.median(ex)), "checked": len(ex)}
        res["grad_rel_err_vs_f64_oracle"] = {"max": float(max(eg)), "median": float(np.median(eg))}
    print(json.dumps(res), flush=True)


def elastic_bench(torch, args):
    """Elastic variant: 1024 QPs n=256, m_eq=32, m_ineq=512 (seed 5), 20% of problems with
    h -= 3|N(0,1)|, penalty 14, float32 fwd (solve + relax) + bwd through solve_qp_elastic_primal
    with loss ||x - 1||^2; one step = forward + backward of the whole batch."""
    import qpax_amd
    from tests import problems as P

    B = args.elastic_batch
    Q, q, A, b, G, h, shifted = P.elastic_infeasible(B, 256, 512, frac=0.2, seed=5, me=32)
    data = [np.asarray(v, np.float32) for v in (Q, q, A, b, G, h)]
    dev = torch.device("cuda", 0)
    ts = [torch.tensor(v, device=dev, requires_grad=True) for v in data]
    pen = 14.0

    def step():
        for t in ts:
            t.grad = None
        x = qpax_amd.solve_qp_elastic_primal(ts[0], ts[1], ts[4], ts[5], pen, A=ts[2], b=ts[3], solver_tol=TOL,
                                             target_kappa=KAPPA)
        ((x - 1.0) ** 2).sum().backward()
        return x

    ms = 1e3 * time_steps(torch, step, args.steps, args.warmup)
    x = step().detach().cpu().numpy().astype(np.float64)
    if args.profile:
        print(json.dumps({"profile_steps": args.steps}), flush=True)
        return
    # the fused solve + relax launch alone, HIP events on the stream the library launches on
    from qpax_amd._tensors import normalize
    from qpax_amd.elastic_qp import _solve

    bt = normalize(*[t.detach() for t in ts])
    pen_t = torch.full((B,), pen, dtype=torch.float32, device=dev)
    st = torch.cuda.current_stream()
    f_ms = []
    for r in range(3):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        outs, conv, iters, routs, rconv, riters = _solve(bt, pen_t, TOL, "fused", relax="reference",
                                                         target_kappa=KAPPA)
        e1.record(st)
        torch.cuda.synchronize()
        if r:
            f_ms.append(e0.elapsed_time(e1))
    f_ms = float(np.mean(f_ms))
    it_f = iters.cpu().numpy().astype(np.float64)
    it_r = riters.cpu().numpy().astype(np.float64)
    fl = sum(flops_per_qp(256, 32, 512, a, b, bwd=False)[0] for a, b in zip(it_f, it_r))
    achieved = fl / (f_ms * 1e-3) / 1e12
    traffic = profiled_traffic("elastic_n256_fp32.md", "k_e_fused_solve")
    res = {
        "metric": "fwd+bwd QPs/s, elastic QP n=256 m_eq=32 m_ineq=512",
        "value": round(B / (ms * 1e-3), 2), "unit": "QPs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{B} elastic QPs n=256 m_eq=32 m_ineq=512, 20% shifted infeasible (h -= 3|N(0,1)|), "
                               "penalty 14, seed 5, float32 inputs/outputs, float64 arithmetic, "
                               "solve_qp_elastic_primal fwd+bwd, tol 1e-5, kappa 1e-3",
                   "batch": B, "io_dtype": "f32"},
        "shifted_frac": float(shifted.mean()),
        "iters_fwd_mean": float(it_f.mean()),
        "iters_relax_mean": float(it_r.mean()),
        "kernel_ms": {"fused_solve_relax": round(f_ms, 3)},
        "roofline": {"kernel": "qpx::k_e_fused_solve<256,float> (solve + relax; float64 MFMA + VALU)",
                     "bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_F64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_F64_TFLOPS, 4), "traffic": traffic,
                     "traffic_source": "profiles/elastic_n256_fp32.md (rocprofv3 FETCH_SIZE x2 + WRITE_SIZE per "
                                       "launch)",
                     **hbm_fields(traffic, f_ms),
                     "algorithmic_flops_per_launch": fl},
    }
    if not args.no_cpu:
        from oracle import qp_oracle as O

        S = args.elastic_check
        d64 = [v[:S].astype(np.float64) for v in data]
        t0 = time.perf_counter()
        xo, rel, info = O.solve_qp_elastic_primal_fwd(d64[0], d64[1], d64[4], d64[5], pen, solver_tol=TOL,
                                                      target_kappa=KAPPA, A=d64[2], b=d64[3])
        cpu_s = time.perf_counter() - t0
        ok = info[0] == 1
        err = np.linalg.norm(x[:S] - xo, axis=1) / np.maximum(np.linalg.norm(xo, axis=1), 1e-12)
        res["x_rel_err_vs_f64_oracle"] = {"max_converged": float(err[ok].max()), "median": float(np.median(err)),
                                          "checked": S}
        res["cpu_baseline"] = {"value": round(S / cpu_s, 3), "unit": "QPs/s (forward only)", "cores": 1,
                               "kind": "port", "sample": f"first {S} problems, forward (solve + relax), float64 "
                                                         "numpy oracle"}
    print(json.dumps(res), flush=True)


def nnls_bench(torch, args):
    """The NNLS batch example's workload (qpax's README and examples/batch_qp.py:14-33): 10000
    nonnegative least-squares QPs, n=5, G=-I, h=0, no equality rows, forward solve_qp only, in
    float64 at solver_tol 1e-10 and in float32 at 1e-5.  One step = one solve of the whole batch;
    `value` is the float64 rate, the float32 one is reported beside it."""
    import qpax_amd
    from tests import problems as P

    B = 10000
    res = {"metric": "forward QPs/s, 10000 NNLS QPs n=5 (G=-I: box rows on the diagonal path)", "unit": "QPs/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "10000 NNLS QPs, F in R^{10x5}, g in R^10 ~ N(0,1), Q=F^T F, q=-F^T g, G=-I5, h=0, "
                                  "seed 1; solve_qp forward", "batch": B}}
    data = P.nnls(B, seed=1)
    for dt, tol, key in ((np.float64, 1e-10, "f64"), (np.float32, 1e-5, "f32")):
        ts = [torch.tensor(np.asarray(v, dt), device="cuda") for v in data]
        out = None

        def step():
            nonlocal out
            out = qpax_amd.solve_qp(*ts, solver_tol=tol)

        ms = 1e3 * time_steps(torch, step, args.steps, args.warmup)
        x = out[0].cpu().numpy().astype(np.float64)
        conv = out[4].cpu().numpy()
        entry = {"qps": round(B / (ms * 1e-3), 1), "ms_per_step": round(ms, 4), "solver_tol": tol,
                 "converged_frac": float((conv == 1).mean()), "iters_mean": float(out[5].float().mean())}
        if not args.no_cpu:
            from oracle import qp_oracle as O

            d64 = [np.asarray(np.asarray(v, dt), np.float64) for v in data]
            xo, so, zo, yo, co, io = O.solve_qp(*d64, solver_tol=tol)
            ok = co == 1
            err = np.linalg.norm(x - xo, axis=1) / np.maximum(np.linalg.norm(xo, axis=1), 1e-12)
            entry["x_rel_err_vs_f64_oracle_max"] = float(err[ok].max())
        res[key] = entry
    res["value"] = res["f64"]["qps"]
    res["ms_per_step"] = res["f64"]["ms_per_step"]
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--cpu-sample", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile", action="store_true", help="only run a few fused steps (for rocprofv3)")
    ap.add_argument("--copy-calib", type=int, default=0, help="MiB of a plain device copy (counter calibration)")
    ap.add_argument("--workload", choices=("c3", "uneven", "elastic", "nnls"), default="c3",
                    help="c3: the headline n=50 B=4096 workload; uneven: 16384 mixed-size QPs (ragged plan); "
                         "elastic: 1024 elastic QPs n=256 m_eq=32 m_ineq=512")
    ap.add_argument("--elastic-batch", type=int, default=1024)
    ap.add_argument("--elastic-check", type=int, default=4, help="problems of the elastic batch checked vs the oracle")
    ap.add_argument("--uneven-batch", type=int, default=16384)
    ap.add_argument("--uneven-check", type=int, default=32, help="problems of the uneven batch checked vs the oracle")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus > 1 is not supported"}))
        sys.exit(2)

    cpu = None
    if args.workload == "c3" and not args.no_cpu and not args.profile and not args.copy_calib:
        # the CPU baseline first, before this process initialises the GPU (its workers are forks)
        from tests import problems as P

        data = [np.asarray(v, np.float32) for v in P.random_feasible(BATCH, N, ME, MI, seed=SEED)]
        cpu = cpu_baseline(data, min(args.cpu_sample, BATCH))

    import torch

    import qpax_amd

    dev = torch.device("cuda", 0)
    if args.copy_calib:
        nb = args.copy_calib << 20
        a = torch.empty(nb // 4, dtype=torch.float32, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calib_bytes": nb}))
        return

    if args.workload == "uneven":
        uneven_bench(torch, args)
        return
    if args.workload == "elastic":
        elastic_bench(torch, args)
        return
    if args.workload == "nnls":
        nnls_bench(torch, args)
        return

    data, ts = make_inputs(torch, dev)
    if args.profile:
        for _ in range(max(args.steps, 1)):
            fwd_bwd(qpax_amd, ts, "fused")
        torch.cuda.synchronize()
        print(json.dumps({"profile_steps": args.steps}), flush=True)
        return

    ms_fused = 1e3 * time_steps(torch, lambda: fwd_bwd(qpax_amd, ts, "fused"), args.steps, args.warmup)
    print(f"# fused fwd+bwd: {ms_fused:.3f} ms/step", file=sys.stderr, flush=True)
    ms_plain = 1e3 * time_steps(torch, lambda: fwd_bwd(qpax_amd, ts, "plain"), args.plain_steps, 1)
    print(f"# plain fwd+bwd: {ms_plain:.3f} ms/step", file=sys.stderr, flush=True)

    f_ms, b_ms, out = kernel_times(torch, qpax_amd, ts, reps=max(3, min(args.steps, 10)))
    it_f = out[5].cpu().numpy().astype(np.float64)
    it_r = out[11].cpu().numpy().astype(np.float64)
    fl_fwd = sum(flops_per_qp(N, ME, MI, a, b, bwd=False)[0] for a, b in zip(it_f, it_r))
    achieved = fl_fwd / (f_ms * 1e-3) / 1e12

    x_gpu = fwd_bwd(qpax_amd, ts, "fused").detach().cpu().numpy()
    g_gpu = [t.grad.detach().cpu().numpy() for t in ts]

    value = BATCH / (ms_fused * 1e-3)
    res = {
        "metric": "fwd+bwd QPs/s @n=50,B=4096: fused vs plain GPU; x/grad err vs FP64",
        "value": round(value, 1),
        "unit": "QPs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_fused, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "solve_qp_primal fwd+bwd, loss ||x-1||^2, 4096 random feasible QPs n=50 m_eq=10 m_ineq=100, "
                        "float32 inputs/outputs, float64 arithmetic, solver_tol=1e-5, target_kappa=1e-3, seed 3",
            "batch": BATCH, "n": N, "m_eq": ME, "m_ineq": MI, "io_dtype": "f32",
        },
        "plain_qps": round(BATCH / (ms_plain * 1e-3), 1),
        "plain_ms_per_step": round(ms_plain, 3),
        "fused_over_plain": round(ms_plain / ms_fused, 2),
        "iters_fwd_mean": float(it_f.mean()),
        "iters_relax_mean": float(it_r.mean()),
        "kernel_ms": {"fused_solve_relax": round(f_ms, 4), "fused_backward": round(b_ms, 4)},
        "roofline": {
            "kernel": "qpx::wave::k_wave_solve<float,64,2,true> (solve + relax; float64 MFMA + VALU)",
            "bound": "mfma",
            "achieved": round(achieved, 3),
            "peak": PEAK_F64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_F64_TFLOPS, 4),
            "traffic": profiled_traffic("k_wave_solve_c3_fp32.md", "k_wave_solve"),
            "traffic_source": "profiles/k_wave_solve_c3_fp32.md (rocprofv3 FETCH_SIZE x2 + WRITE_SIZE per launch)",
            **hbm_fields(profiled_traffic("k_wave_solve_c3_fp32.md", "k_wave_solve"), f_ms),
            "algorithmic_flops_per_launch": fl_fwd,
        },
    }
    if cpu is not None:
        cpu_qps, xo, go, info, host = cpu
        S = xo.shape[0]
        ok = (info[0] == 1) & (info[2] == 1)

        def rel(a, b):
            a = a.reshape(a.shape[0], -1)
            b = b.reshape(b.shape[0], -1)
            return np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-12)

        ex = rel(x_gpu[:S], xo)
        eg = [rel(g[:S], r) for g, r in zip(g_gpu, go)]
        res["x_rel_err_vs_f64_oracle"] = {"max_converged": float(ex[ok].max()), "median": float(np.median(ex)),
                                          "oracle_converged_frac": float(ok.mean()), "sample": S}
        res["grad_rel_err_vs_f64_oracle"] = {k: float(e[ok].max()) for k, e in zip("Q q A b G h".split(), eg)}
        res["cpu_baseline"] = {"value": round(cpu_qps, 2), "unit": "QPs/s", "cores": host["workers"], "kind": "port",
                               "sample": f"{S} of the 4096 QPs, fwd (solve+relax) + bwd, float64 numpy oracle "
                                         f"(QR solves), one single-threaded process per core",
                               **host}
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
