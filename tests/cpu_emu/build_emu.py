"""Host (x86-64) build of libqpx's HIP sources against the CPU emulator, with AddressSanitizer.

Produces tests/cpu_emu/_build/libqpx_emu.so: the library's C ABI (include/qpx.h) with every
kernel running on the CPU emulator of hip/hip_runtime.h, plus emu_alloc/emu_free for exact-size
buffers.  Load it only in a process started with the AddressSanitizer runtime preloaded
(asan_runtime() gives its path).  Rebuilds a unit only when a source or header is newer than its
object.

Usage: python tests/cpu_emu/build_emu.py
"""

from __future__ import annotations

import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
CSRC = os.path.join(ROOT, "qpax_amd", "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libqpx_emu.so")
CLANG = "/opt/rocm/llvm/bin/clang++"

FLAGS = ["-x", "c++", "-std=c++17", "-O1", "-g", "-fPIC", "-fsanitize=address", "-fno-omit-frame-pointer",
         "-ffp-contract=off", "-Wno-unused-command-line-argument", "-I", HERE, "-I", os.path.join(ROOT, "include")]

sys.path.insert(0, os.path.join(ROOT, "qpax_amd"))
import build as _hipbuild  # noqa: E402  (the instance lists of the GPU build)


def asan_runtime() -> str:
    hits = glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so")
    if not hits:
        raise RuntimeError("AddressSanitizer runtime of /opt/rocm/llvm not found")
    return hits[0]


def _units():
    units = [(os.path.join(CSRC, "qpx_solver.hip"), "qpx_solver.o", []),
             (os.path.join(CSRC, "qpx_ragged.hip"), "qpx_ragged.o", []),
             (os.path.join(HERE, "emu_rt.cpp"), "emu_rt.o", [])]
    for ts, nmax, mic in _hipbuild.WAVE_INSTANCES:
        units.append((os.path.join(CSRC, "qpx_wave_inst.hip"), f"wave_{ts[0]}_{nmax}_{mic}.o",
                      [f"-DQPX_TS={ts}", f"-DQPX_NMAX={nmax}", f"-DQPX_MIC={mic}"]))
    for ts, nti, w in _hipbuild.GROUP_INSTANCES:
        units.append((os.path.join(CSRC, "qpx_group_inst.hip"), f"group_{ts[0]}_{nti}_{w}.o",
                      [f"-DQPX_TS={ts}", f"-DQPX_NTI={nti}", f"-DQPX_W={w}"]))
    return units


def _newest_dep() -> float:
    deps = glob.glob(os.path.join(CSRC, "*")) + glob.glob(os.path.join(ROOT, "include", "*.h"))
    deps += [os.path.join(HERE, "hip", "hip_runtime.h"), os.path.join(HERE, "emu_rt.cpp"), __file__]
    return max(os.path.getmtime(p) for p in deps)


def _compile(unit, newest):
    src, obj, defs = unit
    obj = os.path.join(OUT, obj)
    if os.path.exists(obj) and os.path.getmtime(obj) >= newest:
        return None
    cmd = [CLANG, *FLAGS, *defs, "-c", src, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        return f"{' '.join(cmd)}\n{res.stdout}{res.stderr}"
    return None


def build(jobs: int | None = None) -> str:
    os.makedirs(OUT, exist_ok=True)
    newest = _newest_dep()
    units = _units()
    jobs = jobs or max(1, min(8, os.cpu_count() or 2))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        errs = [e for e in ex.map(lambda u: _compile(u, newest), units) if e]
    if errs:
        sys.stderr.write("\n".join(errs[:3]))
        raise RuntimeError(f"host build of {len(errs)} unit(s) failed")
    sw = os.path.join(OUT, "emu_switch.o")
    if not os.path.exists(sw) or os.path.getmtime(sw) < os.path.getmtime(os.path.join(HERE, "emu_switch.S")):
        subprocess.run([CLANG, "-c", os.path.join(HERE, "emu_switch.S"), "-o", sw], check=True)
    objs = [os.path.join(OUT, u[1]) for u in units] + [sw]
    if os.path.exists(LIB) and os.path.getmtime(LIB) >= max(os.path.getmtime(o) for o in objs):
        return LIB
    tmp = LIB + ".tmp"
    subprocess.run([CLANG, "-shared", "-fsanitize=address", "-shared-libasan", *objs, "-o", tmp], check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
