"""Builds the HIP library libcardioid_hip.so in-tree for gfx950.

    python -m cardioid_amd.build        (or __graft_entry__.build())

Every HIP source is compiled with hipcc for gfx950 only; the FP64 kernels of
the drop-in path with -ffp-contract=off (no fused multiply-add, the
reference's rounding); packed FP32 instructions are disabled throughout.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcardioid_hip.so")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-fopenmp",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument"]
# per-source extra flags
SOURCES = {
    "kernels_f64.hip": ["-ffp-contract=off"],
    "engine_f64.hip": ["-ffp-contract=off"],
    "engine_mixed.hip": ["-ffp-contract=fast"],
    "engine_host.hip": ["-ffp-contract=off"],
    "slab.cpp": ["-x", "hip", "-ffp-contract=off"],
    "plugins.hip": ["-ffp-contract=off"],
    "sim.hip": ["-ffp-contract=off"],
    "batch.hip": ["-ffp-contract=off"],
    "setup.cpp": ["-x", "hip", "-ffp-contract=off"],
    "fgr_weights.cpp": ["-x", "hip", "-ffp-contract=off"],
    "tt06_host.cpp": ["-x", "hip", "-ffp-contract=off"],
    "host/reaction.cpp": ["-x", "hip", "-ffp-contract=off"],
    "host/diffusion.cpp": ["-x", "hip", "-ffp-contract=off"],
    "host/simulate.hip": ["-ffp-contract=off"],
    "checkpoint.cpp": ["-x", "hip", "-ffp-contract=off"],
    "setup_device.hip": ["-ffp-contract=off"],
    "stream_probe.hip": ["-ffp-contract=off"],
}


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src: str, extra):
    obj = os.path.join(BUILD, os.path.splitext(src.replace("/", "_"))[0] + ".o")
    cmd = [hipcc(), *COMMON, *extra, "-fgpu-rdc", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in (r.stderr or "").splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{err}")
    return obj, err


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        results = list(ex.map(lambda kv: _compile(*kv), SOURCES.items()))
    for _, err in results:
        if err.strip() and verbose:
            print(err, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-fgpu-rdc", "--hip-link", "-shared", "-fopenmp",
           "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
