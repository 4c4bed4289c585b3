"""Build the in-tree native library ``within_amd/_lib/libwithin_amd.so``.

* ``csrc/plan.cpp`` (host setup, OpenMP) with g++;
* ``csrc/kernels.hip``, ``csrc/multi.hip``, ``csrc/gmres.hip``, ``csrc/coarse.hip``, ``csrc/dense.hip`` and ``csrc/runtime.cpp`` (device
  code + C ABI) with hipcc for gfx950, packed FP32 instructions disabled
  (``dense.hip`` also without FP contraction);
* linked into one shared library exporting the symbols of
  ``include/within_amd.h``.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libwithin_amd.so")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(line for line in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in line and "unknown target feature" not in line)
    if r.returncode != 0:
        raise RuntimeError(f"command failed: {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)


def build(verbose: bool = False) -> str:
    """Compile every source and link the library; returns its path."""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    jobs = []
    plan_o = os.path.join(BUILD, "plan.o")
    # no FP contraction: host factors must match the oracle bit for bit
    jobs.append((["g++", "-O3", "-ffp-contract=off", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-c",
                  os.path.join(CSRC, "plan.cpp"), "-o", plan_o], plan_o))
    for src in ("kernels.hip", "multi.hip", "gmres.hip", "coarse.hip", "runtime.cpp", "dense.hip"):
        o = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        # the dense-core inverse must round like the host factorisation
        exact = ["-ffp-contract=off"] if src == "dense.hip" else []
        jobs.append(([hipcc, *HIP_FLAGS, *exact, *lang, "-c", os.path.join(CSRC, src), "-o", o], o))
    # the translation units are independent: compile them concurrently
    with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        for f in [ex.submit(_run, cmd, verbose) for cmd, _ in jobs]:
            f.result()
    objs = [o for _, o in jobs]
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lgomp"], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
