"""Build libqmetal.so in-tree: hipcc for the gfx950 kernels and executor,
g++ for the plain C++ core (plan builder, gate matrices, C ABI).

    python -m qmetal_amd.build          # or qmetal_amd.build.build()

The plan builder is compiled with -ffp-contract=off so its FP64 matrix
products round exactly like the reference builder's.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libqmetal.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("QMETAL_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["kernels.hip", "simulator.hip", "batch.hip"]
# low_kernels.hip is built once per pair of step counts (its kernel instances
# take most of the build; the objects compile in parallel).
LOW_PAIRS = [5, 7, 9, 11]
CXX_SOURCES = ["gates.cpp", "plan.cpp", "schedule.cpp", "capi.cpp"]

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # Packed FP32 instructions stay off on this platform; hipcc -O3 would
    # otherwise form v_pk_* from ordinary float code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=fast",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra"]


def _run(cmd):
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + proc.stdout + proc.stderr)
    return proc


def build(verbose: bool = False) -> str:
    os.makedirs(OUT, exist_ok=True)
    # The pass kernel's slot statements are generated text.
    _run([sys.executable, os.path.join(CSRC, "gen_slot_asm.py")])
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OUT, src + ".o")
        jobs.append((obj, [os.path.join(ROCM, "bin", "hipcc"), *HIPCC_FLAGS, *inc, "-c",
                           os.path.join(CSRC, src), "-o", obj]))
    for s0 in LOW_PAIRS:
        obj = os.path.join(OUT, f"low_kernels.{s0}.o")
        jobs.append((obj, [os.path.join(ROCM, "bin", "hipcc"), *HIPCC_FLAGS, *inc,
                           f"-DQM_LOW_S0={s0}", "-c", os.path.join(CSRC, "low_kernels.hip"),
                           "-o", obj]))
    for src in CXX_SOURCES:
        obj = os.path.join(OUT, src + ".o")
        jobs.append((obj, ["g++", *CXX_FLAGS, *inc, "-c", os.path.join(CSRC, src), "-o", obj]))
    # Longest first: the low kernel objects, then the rest.
    jobs.sort(key=lambda j: "low_kernels" not in j[0])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for (obj, cmd), _ in zip(jobs, ex.map(lambda j: _run(j[1]), jobs)):
            if verbose:
                print(" ".join(cmd))
    tmp = LIB + ".tmp"
    _run([os.path.join(ROCM, "bin", "hipcc"), f"--offload-arch={ARCH}", "-shared", "-o", tmp,
          *[o for o, _ in jobs], f"-L{ROCM}/lib", "-lamdhip64", f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
