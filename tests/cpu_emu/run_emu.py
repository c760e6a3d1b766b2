"""Runs libqpx entry points on the CPU emulator (tests/cpu_emu/_build/libqpx_emu.so).

Started by the tests as a child process with the AddressSanitizer runtime preloaded:

    LD_PRELOAD=<asan runtime> python tests/cpu_emu/run_emu.py <case.npz> <out.npz>

Every device buffer is an exact-size allocation, so an out-of-bounds access by any kernel (global
memory or the workgroup's LDS image) ends the process with an AddressSanitizer report and a
non-zero exit status.  Case kinds (the npz's `kind`):

* "ragged": problems given as flat arrays Q, q, A, b, G, h (problem after problem) with sizes n,
  me, mi; runs qpx_plan_create, qpx_plan_solve_relaxed and qpx_plan_diff with dl/dx = 2 (x - 1);
  optional `tune` = 1 runs qpx_plan_tune (every candidate launch shape) for the forward.
* "uniform": batched arrays Q (B,n,n), ...; runs qpx_solve_qp_relaxed and qpx_diff_qp in `mode`.
* "init", "init_elastic", "trace": qpx_initialize, qpx_initialize_elastic (with `penalty`) and
  qpx_solve_qp_trace (with `tol`, `debug_step`) on batched arrays.
* "elastic": qpx_solve_qp_elastic_relaxed and qpx_diff_qp_elastic (with `penalty`, `mode`).

The output npz holds every output array (flat, in the layout of the C ABI).
"""

from __future__ import annotations

import ctypes
import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
LIB = os.path.join(HERE, "_build", "libqpx_emu.so")


def _load_signatures():
    spec = importlib.util.spec_from_file_location("_qpx_lib_sigs", os.path.join(ROOT, "qpax_amd", "_lib.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


class Emu:
    def __init__(self):
        self.sig = _load_signatures()
        lib = ctypes.CDLL(LIB)
        for name, (res, args) in self.sig.SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        lib.emu_alloc.restype = ctypes.c_void_p
        lib.emu_alloc.argtypes = [ctypes.c_size_t]
        lib.emu_free.argtypes = [ctypes.c_void_p]
        self.lib = lib
        self.bufs = []

    def dev(self, arr):
        """Exact-size device copy of a numpy array (None for an empty array)."""
        arr = np.ascontiguousarray(arr)
        if arr.size == 0:
            return None
        p = self.lib.emu_alloc(arr.nbytes)
        ctypes.memmove(p, arr.ctypes.data, arr.nbytes)
        self.bufs.append(p)
        return p

    def empty(self, count, dtype):
        if count == 0:
            return None
        p = self.lib.emu_alloc(int(count) * np.dtype(dtype).itemsize)
        ctypes.memset(p, 0xFF, int(count) * np.dtype(dtype).itemsize)
        self.bufs.append(p)
        return p

    @staticmethod
    def host(p, count, dtype):
        out = np.empty(int(count), dtype=dtype)
        if count:
            ctypes.memmove(out.ctypes.data, p, out.nbytes)
        return out

    def check(self, st, what):
        if st != 0:
            raise RuntimeError(f"{what}: status {st}: {self.lib.qpx_last_error().decode()}")


def run_ragged(e: Emu, c) -> dict:
    lib = e.lib
    dt = np.float32 if str(c["dtype"]) == "float32" else np.float64
    code = 0 if dt == np.float32 else 1
    n, me, mi = (np.ascontiguousarray(c[k], dtype=np.int32) for k in ("n", "me", "mi"))
    B = len(n)
    plan = ctypes.c_void_p()
    e.check(lib.qpx_plan_create(code, B, n.ctypes.data, me.ctypes.data, mi.ctypes.data, ctypes.byref(plan)),
            "plan_create")
    info = np.zeros(8, np.int64)
    e.check(lib.qpx_plan_info(plan, info.ctypes.data), "plan_info")
    tot = info[:6]
    ops = [e.dev(np.asarray(c[k], dt)) for k in ("Q", "q", "A", "b", "G", "h")]
    ws = e.empty(info[6], np.uint8)
    x, s, z, y = (e.empty(tot[k], dt) for k in (1, 5, 5, 3))
    conv, it = e.empty(B, np.int32), e.empty(B, np.int32)
    xr, sr, zr, yr = (e.empty(tot[k], dt) for k in (1, 5, 5, 3))
    rconv, rit = e.empty(B, np.int32), e.empty(B, np.int32)
    # with `tune`, qpx_plan_tune runs every candidate launch shape of every bucket (wavefronts per
    # workgroup 1/2/4, workgroup sizes, persistent or not) before the final call
    fn = lib.qpx_plan_tune if ("tune" in c.files and int(c["tune"])) else lib.qpx_plan_solve_relaxed
    st = fn(plan, *ops, float(c["tol"]), float(c["kappa"]), x, s, z, y, conv, it, xr, sr, zr, yr, rconv, rit, ws,
            int(info[6]), None)
    e.check(st, "plan_solve_relaxed")
    xh = e.host(x, tot[1], dt)
    g = e.dev((2.0 * (xh - 1.0)).astype(dt))
    grads = [e.empty(tot[k], dt) for k in range(6)]
    st = lib.qpx_plan_diff(plan, *ops, xr, sr, zr, yr, g, 0, *grads, ws, int(info[6]), None)
    e.check(st, "plan_diff")
    out = {"x": xh, "s": e.host(s, tot[5], dt), "z": e.host(z, tot[5], dt), "y": e.host(y, tot[3], dt),
           "conv": e.host(conv, B, np.int32), "iters": e.host(it, B, np.int32),
           "xr": e.host(xr, tot[1], dt), "sr": e.host(sr, tot[5], dt), "zr": e.host(zr, tot[5], dt),
           "yr": e.host(yr, tot[3], dt), "rconv": e.host(rconv, B, np.int32), "riters": e.host(rit, B, np.int32)}
    for k, name in enumerate(("dQ", "dq", "dA", "db", "dG", "dh")):
        out[name] = e.host(grads[k], tot[k], dt)
    buckets = []
    for i in range(int(info[7])):
        o = np.zeros(10, np.int64)
        e.check(lib.qpx_plan_bucket(plan, i, o.ctypes.data), "plan_bucket")
        buckets.append(o)
    out["buckets"] = np.array(buckets)
    off = np.zeros((B, 6), np.int64)
    e.check(lib.qpx_plan_offsets(plan, off.ctypes.data), "plan_offsets")
    out["offsets"] = off
    lib.qpx_plan_destroy(plan)
    return out


def run_uniform(e: Emu, c) -> dict:
    lib = e.lib
    dt = np.float32 if str(c["dtype"]) == "float32" else np.float64
    code = 0 if dt == np.float32 else 1
    Q, q, A, b, G, h = (np.asarray(c[k], dt) for k in ("Q", "q", "A", "b", "G", "h"))
    B, n, me, mi = Q.shape[0], Q.shape[1], A.shape[1], G.shape[1]
    mode = 1 if str(c["mode"]) == "fused" else 0
    ops = [e.dev(v) for v in (Q, q, A, b, G, h)]
    wsb = lib.qpx_workspace_size(B, n, me, mi)
    ws = e.empty(wsb, np.uint8)
    x, s, z, y = e.empty(B * n, dt), e.empty(B * mi, dt), e.empty(B * mi, dt), e.empty(B * me, dt)
    xr, sr, zr, yr = e.empty(B * n, dt), e.empty(B * mi, dt), e.empty(B * mi, dt), e.empty(B * me, dt)
    conv, it, rconv, rit = (e.empty(B, np.int32) for _ in range(4))
    st = lib.qpx_solve_qp_relaxed(code, mode, B, n, me, mi, *ops, float(c["tol"]), float(c["kappa"]), x, s, z, y,
                                  conv, it, xr, sr, zr, yr, rconv, rit, ws, wsb, None)
    e.check(st, "solve_qp_relaxed")
    xh = e.host(x, B * n, dt)
    g = e.dev((2.0 * (xh - 1.0)).astype(dt))
    sizes = (B * n * n, B * n, B * me * n, B * me, B * mi * n, B * mi)
    grads = [e.empty(k, dt) for k in sizes]
    st = lib.qpx_diff_qp(code, mode, B, n, me, mi, *ops, xr, sr, zr, yr, g, 0, *grads, ws, wsb, None)
    e.check(st, "diff_qp")
    out = {"x": xh, "s": e.host(s, B * mi, dt), "z": e.host(z, B * mi, dt), "y": e.host(y, B * me, dt),
           "conv": e.host(conv, B, np.int32), "iters": e.host(it, B, np.int32),
           "xr": e.host(xr, B * n, dt), "sr": e.host(sr, B * mi, dt), "zr": e.host(zr, B * mi, dt),
           "yr": e.host(yr, B * me, dt), "rconv": e.host(rconv, B, np.int32), "riters": e.host(rit, B, np.int32)}
    for k, name in enumerate(("dQ", "dq", "dA", "db", "dG", "dh")):
        out[name] = e.host(grads[k], sizes[k], dt)
    return out


def _batch_args(e: Emu, c, names):
    dt = np.float32 if str(c["dtype"]) == "float32" else np.float64
    arrs = [np.asarray(c[k], dt) for k in names]
    return dt, (0 if dt == np.float32 else 1), arrs


def run_init(e: Emu, c) -> dict:
    """qpx_initialize on batched arrays Q, q, A, b, G, h."""
    lib = e.lib
    dt, code, (Q, q, A, b, G, h) = _batch_args(e, c, ("Q", "q", "A", "b", "G", "h"))
    B, n, me, mi = Q.shape[0], Q.shape[1], A.shape[1], G.shape[1]
    ops = [e.dev(v) for v in (Q, q, A, b, G, h)]
    wsb = lib.qpx_workspace_size(B, n, me, mi)
    ws = e.empty(wsb, np.uint8)
    x, s, z, y = e.empty(B * n, dt), e.empty(B * mi, dt), e.empty(B * mi, dt), e.empty(B * me, dt)
    e.check(lib.qpx_initialize(code, B, n, me, mi, *ops, x, s, z, y, ws, wsb, None), "initialize")
    return {"x": e.host(x, B * n, dt), "s": e.host(s, B * mi, dt), "z": e.host(z, B * mi, dt),
            "y": e.host(y, B * me, dt)}


def run_init_elastic(e: Emu, c) -> dict:
    """qpx_initialize_elastic on batched arrays Q, q, A, b, G, h and penalty (B,)."""
    lib = e.lib
    dt, code, (Q, q, A, b, G, h, pen) = _batch_args(e, c, ("Q", "q", "A", "b", "G", "h", "penalty"))
    B, n, me, mi = Q.shape[0], Q.shape[1], A.shape[1], G.shape[1]
    ops = [e.dev(v) for v in (Q, q, A, b, G, h, pen)]
    wsb = lib.qpx_workspace_size(B, n, me, mi)
    ws = e.empty(wsb, np.uint8)
    outs = [e.empty(B * n, dt)] + [e.empty(B * mi, dt) for _ in range(5)] + [e.empty(B * me, dt)]
    e.check(lib.qpx_initialize_elastic(code, B, n, me, mi, *ops, *outs, ws, wsb, None), "initialize_elastic")
    sizes = [B * n] + [B * mi] * 5 + [B * me]
    return {k: e.host(o, sz, dt) for k, o, sz in zip(("x", "t", "s1", "s2", "z1", "z2", "y"), outs, sizes)}


def run_trace(e: Emu, c) -> dict:
    """qpx_solve_qp_trace on batched arrays; `debug_step` selects QPX_TRACE_DEBUG_STEP."""
    lib = e.lib
    dt, code, (Q, q, A, b, G, h) = _batch_args(e, c, ("Q", "q", "A", "b", "G", "h"))
    B, n, me, mi = Q.shape[0], Q.shape[1], A.shape[1], G.shape[1]
    ops = [e.dev(v) for v in (Q, q, A, b, G, h)]
    wsb = lib.qpx_workspace_size(B, n, me, mi)
    ws = e.empty(wsb, np.uint8)
    x, s, z, y = e.empty(B * n, dt), e.empty(B * mi, dt), e.empty(B * mi, dt), e.empty(B * me, dt)
    conv, it = e.empty(B, np.int32), e.empty(B, np.int32)
    tr = e.empty(B * 30 * 5, np.float64)
    flags = 4 if int(c["debug_step"]) else 0
    e.check(lib.qpx_solve_qp_trace(code, B, n, me, mi, *ops, float(c["tol"]), flags, x, s, z, y, conv, it, tr, ws,
                                   wsb, None), "solve_qp_trace")
    return {"x": e.host(x, B * n, dt), "s": e.host(s, B * mi, dt), "z": e.host(z, B * mi, dt),
            "y": e.host(y, B * me, dt), "conv": e.host(conv, B, np.int32), "iters": e.host(it, B, np.int32),
            "trace": e.host(tr, B * 150, np.float64).reshape(B, 30, 5)}


def run_elastic(e: Emu, c) -> dict:
    """qpx_solve_qp_elastic_relaxed (relaxation `flags`: 0 the reference's, 2 consistent signs) and
    qpx_diff_qp_elastic with dl/dx = 2 (x - 1) on batched arrays Q, q, A, b, G, h and penalty (B,),
    in `mode`."""
    lib = e.lib
    dt, code, (Q, q, A, b, G, h, pen) = _batch_args(e, c, ("Q", "q", "A", "b", "G", "h", "penalty"))
    B, n, me, mi = Q.shape[0], Q.shape[1], A.shape[1], G.shape[1]
    mode = 1 if str(c["mode"]) == "fused" else 0
    ops = [e.dev(v) for v in (Q, q, A, b, G, h, pen)]
    wsb = lib.qpx_workspace_size(B, n, me, mi)
    ws = e.empty(wsb, np.uint8)
    sizes = [B * n] + [B * mi] * 5 + [B * me]
    outs = [e.empty(k, dt) for k in sizes]
    routs = [e.empty(k, dt) for k in sizes]
    conv, it, rconv, rit = (e.empty(B, np.int32) for _ in range(4))
    flags = int(c["flags"]) if "flags" in c.files else 0
    e.check(lib.qpx_solve_qp_elastic_relaxed(code, mode, B, n, me, mi, *ops, float(c["tol"]), float(c["kappa"]), flags,
                                             *outs, conv, it, *routs, rconv, rit, ws, wsb, None),
            "solve_qp_elastic_relaxed")
    xh = e.host(outs[0], B * n, dt)
    g = e.dev((2.0 * (xh - 1.0)).astype(dt))
    gsz = (B * n * n, B * n, B * me * n, B * me, B * mi * n, B * mi)
    grads = [e.empty(k, dt) for k in gsz]
    e.check(lib.qpx_diff_qp_elastic(code, mode, B, n, me, mi, *ops, *routs, g, 0, *grads, ws, wsb, None),
            "diff_qp_elastic")
    names = ("x", "t", "s1", "s2", "z1", "z2", "y")
    out = {k: e.host(o, sz, dt) for k, o, sz in zip(names, outs, sizes)}
    out.update({k + "r": e.host(o, sz, dt) for k, o, sz in zip(names, routs, sizes)})
    out.update({"conv": e.host(conv, B, np.int32), "iters": e.host(it, B, np.int32),
                "rconv": e.host(rconv, B, np.int32), "riters": e.host(rit, B, np.int32)})
    for k, name in enumerate(("dQ", "dq", "dA", "db", "dG", "dh")):
        out[name] = e.host(grads[k], gsz[k], dt)
    return out


def main(argv):
    case = np.load(argv[1], allow_pickle=False)
    e = Emu()
    kind = str(case["kind"])
    out = {"ragged": run_ragged, "uniform": run_uniform, "init": run_init, "init_elastic": run_init_elastic,
           "trace": run_trace, "elastic": run_elastic}[kind](e, case)
    np.savez(argv[2], **out)
    print(f"emu: {kind} case done", flush=True)


if __name__ == "__main__":
    main(sys.argv)
