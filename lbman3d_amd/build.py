"""Build lbman3d_amd/liblbman3d.so from the HIP sources in lbman3d_amd/csrc
with hipcc for gfx950 (objects in lbman3d_amd/_build/, library in-tree)."""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblbman3d.so")
ARCH = os.environ.get("LBM_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++17", "-Wall",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", os.path.join(HERE, "..", "include")]
# Per-source flags: the FP64 step and the host runtime keep the reference's
# operand order without FMA contraction (bitwise parity); FP32 kernels contract.
SOURCES = {
    "capi.cpp": ["-x", "hip", "-ffp-contract=off"],
    "kernels_fp64.hip": ["-ffp-contract=off"],
    "kernels_fp32.hip": ["-ffp-contract=fast", "-munsafe-fp-atomics"],
    "batch.hip": ["-ffp-contract=fast", "-x", "hip"],
    "closure.hip": ["-ffp-contract=fast"],
    "defects.hip": ["-ffp-contract=off"],
}
EXTRA: list[str] = []


def _hipcc() -> str:
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c
    raise RuntimeError("hipcc not found")


def _compile(src: str, flags: list[str]) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [_hipcc(), *COMMON, *flags, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(line for line in r.stderr.splitlines() if "packed-fp32-ops" not in line
                     and "warning: feature flag" not in line and line.strip())
    if warn:
        sys.stderr.write(warn + "\n")
    return obj


def build(jobs: int = 4) -> str:
    """Compile every HIP source for gfx950 and link liblbman3d.so. Always
    recompiles (this is the build check)."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = dict(SOURCES)
    for e in EXTRA:
        if os.path.exists(os.path.join(CSRC, e)):
            srcs[e] = ["-ffp-contract=fast"]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda kv: _compile(*kv), srcs.items()))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


HOST_CHECK_SRC = os.path.join(HERE, "..", "tests", "cpp", "host_api_check.cpp")
HOST_CHECK_BIN = os.path.join(HERE, "..", "tests", "cpp", "host_api_check")


def build_host_check() -> str:
    """Compile the C++ host-interface check (tests/cpp/host_api_check.cpp)
    against include/lbman3d.hpp and the in-tree library."""
    cmd = ["g++", "-std=c++20", "-O1", "-Wall", "-I", os.path.join(HERE, "..", "include"), HOST_CHECK_SRC,
           "-L", HERE, "-llbman3d", "-Wl,-rpath,$ORIGIN/../../lbman3d_amd", "-o", HOST_CHECK_BIN]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host check build failed:\n{' '.join(cmd)}\n{r.stderr}")
    return HOST_CHECK_BIN


if __name__ == "__main__":
    print(build())
    print(build_host_check())
