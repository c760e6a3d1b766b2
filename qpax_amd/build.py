"""Build libqpx.so (HIP, gfx950) in-tree with hipcc.

Every wavefront-kernel instantiation is its own object (they are large, fully unrolled register
kernels), compiled in parallel, then linked with the generic solver and the C ABI.

Usage: python -m qpax_amd.build   (or qpax_amd.build.build() from Python)
"""

from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libqpx.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 (v_pk_*_f32) generation is switched off for every HIP source.
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA accumulators in ordinary VGPRs: avoids AGPR<->VGPR copies around every MFMA when the
    # accumulators are loop-carried next to a large register-resident working set
    "-mllvm", "-amdgpu-mfma-vgpr-form=true",
    "-Wno-unused-command-line-argument",
]

# (storage type, NMAX, MIC) instantiations of the wavefront-per-QP kernels
WAVE_INSTANCES = [(ts, nmax, mic) for ts in ("float", "double")
                  for (nmax, mic) in ((32, 1), (32, 2), (64, 1), (64, 2), (64, 4))]


# (storage type, NTI, W) instantiations of the workgroup-per-QP matrix-core kernels
GROUP_INSTANCES = [(ts, nti, w) for ts in ("float", "double") for nti in range(4, 11) for w in (4, 8)]


def _units():
    inc = ["-I", os.path.join(HERE, "..", "include")]
    units = [(os.path.join(CSRC, "qpx_solver.hip"), os.path.join(OBJ, "qpx_solver.o"), inc),
             (os.path.join(CSRC, "qpx_ragged.hip"), os.path.join(OBJ, "qpx_ragged.o"), inc)]
    for ts, nmax, mic in WAVE_INSTANCES:
        defs = [f"-DQPX_TS={ts}", f"-DQPX_NMAX={nmax}", f"-DQPX_MIC={mic}"]
        units.append((os.path.join(CSRC, "qpx_wave_inst.hip"),
                      os.path.join(OBJ, f"qpx_wave_{ts[0]}_{nmax}_{mic}.o"), inc + defs))
    for ts, nti, w in GROUP_INSTANCES:
        defs = [f"-DQPX_TS={ts}", f"-DQPX_NTI={nti}", f"-DQPX_W={w}"]
        units.append((os.path.join(CSRC, "qpx_group_inst.hip"),
                      os.path.join(OBJ, f"qpx_group_{ts[0]}_{nti}_{w}.o"), inc + defs))
    return units


def _compile(unit, verbose):
    src, obj, extra = unit
    cmd = [HIPCC, *FLAGS, *extra, "-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        return f"{' '.join(cmd)}\n{res.stdout}{res.stderr}"
    return None


def build(verbose: bool = False, jobs: int | None = None) -> str:
    os.makedirs(OBJ, exist_ok=True)
    units = _units()
    jobs = jobs or min(len(units), max(1, (os.cpu_count() or 2)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        errs = [e for e in ex.map(lambda u: _compile(u, verbose), units) if e]
    if errs:
        sys.stderr.write("\n".join(errs))
        raise RuntimeError(f"hipcc failed building {len(errs)} unit(s) of {LIB}")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *[u[1] for u in units], "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"link of {LIB} failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
