#!/usr/bin/env python3
"""Throughput of Cardioid's electrics step on one MI355X.

Workload (default): 256x256x512 grid (33.5M points) with a truncated
ellipsoidal-shell anatomy at ~60% tissue (endo/mid/epi by transmural depth),
fibre conductivity with the helix angle rotating -60..+60 deg epi->endo,
8 seeded endocardial stimuli, dt = 0.02 ms, plain FP64 (the drop-in
arithmetic), one fused launch per step.  A "step" is one omp-loop step over
every tissue cell (stimulus, TT06Dev reaction, FGR stencil, Vm update,
activation sensors).

Prints one JSON line (rank 0).  value = tissue cell-steps per second of the
fused FP64 engine (Gcell-steps/s), measured with HIP events on the engine's
stream over exactly --steps steps after --warmup untimed ones.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gcell-steps/s at 33M cells, FP64 vs mixed-FP32; p50/p99 batched-request latency"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per tissue cell-step of the fused FP64 step: 19 FP64
# states read+written (304), Vm read+written (16), 18 float weights + a double
# centre weight (80)
FUSED_F64_BYTES = 400
UNFUSED_F64_BYTES = 480
# mixed: 17 FP32 states r+w (136), Ca_i/Na_i FP64 r+w (32), Vm r+w (16), 18 FP32 weights (72)
MIXED_BYTES = 256


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(budget_s: float = 12.0):
    """The oracle (a C/OpenMP restatement of the reference CPU path) on a
    bounded sample of the same workload: a 96x96x192 ellipsoidal shell from
    the same generator, as many steps as fit in `budget_s`."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as OL
    from cardioid_amd import workloads as W
    threads = min(16, os.cpu_count() or 1)
    a, depth = W.ellipsoid_shell(96, 96, 192, seed=0)
    stims = W.endocardial_point_stimuli(a, depth, 8, seed=0)
    sim = OL.OracleSim(a, W.CELLTYPES_EMI, 0.02, W.DIFFUSION_SCALE, stims, nthreads=threads)
    sim.run(2)
    steps = 0
    t0 = time.perf_counter()
    while True:
        sim.run(5)
        steps += 5
        el = time.perf_counter() - t0
        if el > budget_s:
            break
    sim.close()
    rate = a.ncell * steps / el / 1e9
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": rate, "unit": "Gcell-steps/s", "cores": threads, "kind": "port",
            "sample": f"96x96x192 ellipsoidal shell ({a.ncell} tissue cells), {steps} steps, "
                      f"{el:.1f} s, oracle/liboracle.so with {threads} OpenMP threads on {cpu}"}


def load_traffic():
    path = os.path.join(ROOT, "profiles", "step_fused_f64_shell256x256x512_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get("bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--grid", type=str, default="256,256,512")
    ap.add_argument("--unfused-steps", type=int, default=500)
    ap.add_argument("--mixed-steps", type=int, default=5000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: every box in this setup has a single GPU and the "
              "path is not sharded across GPUs here (--gpus must be 1).", file=sys.stderr)
        sys.exit(2)

    from cardioid_amd import hip, workloads as W
    nx, ny, nz = (int(v) for v in args.grid.split(","))
    log(f"generating {nx}x{ny}x{nz} ellipsoidal shell")
    t = time.perf_counter()
    anat, depth = W.ellipsoid_shell(nx, ny, nz, seed=0)
    stims = W.endocardial_point_stimuli(anat, depth, 8, seed=0)
    log(f"{anat.ncell} tissue cells ({anat.ncell / (nx * ny * nz):.3f} of the grid) "
        f"in {time.perf_counter() - t:.1f} s")

    results = {}
    for mode, steps in (("fp64_fused", args.steps), ("mixed_fused", args.mixed_steps),
                        ("fp64_unfused", args.unfused_steps)):
        if steps <= 0:
            continue
        t = time.perf_counter()
        sim = hip.Simulation(anat, W.CELLTYPES_EMI, 0.02, W.DIFFUSION_SCALE, stims, mode=mode)
        log(f"{mode}: engine ready in {time.perf_counter() - t:.1f} s, {sim.info()}")
        sim.run(args.warmup)
        sim.sync()
        ms = sim.run_timed(steps)
        vm = sim.vm()
        act = sim.activation()[1]
        sim.close()
        rate = anat.ncell * steps / (ms * 1e-3) / 1e9
        log(f"{mode}: {steps} steps in {ms:.1f} ms -> {rate:.3f} Gcell-steps/s; "
            f"Vm range [{vm.min():.2f}, {vm.max():.2f}], activated {(act >= 0).mean():.3f}")
        results[mode] = dict(ms=ms, steps=steps, rate=rate)

    fused = results["fp64_fused"]
    ms_per_step = fused["ms"] / fused["steps"]
    achieved = FUSED_F64_BYTES * anat.ncell / (ms_per_step * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": fused["rate"],
        "unit": "Gcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"ellipsoidal shell {nx}x{ny}x{nz}, ~60% tissue, endo/mid/epi, "
                               "rotating fibres, 8 endocardial stimuli, dt 0.02 ms, plain FP64, "
                               "fused step",
                   "grid_points": nx * ny * nz, "tissue_cells": anat.ncell, "dt_ms": 0.02,
                   "mode": "fp64_fused"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic(),
                     "kernel": "k_step_fused_f64", "algorithmic_bytes_per_cell_step": FUSED_F64_BYTES},
    }
    if "mixed_fused" in results:
        mx = results["mixed_fused"]
        mms = mx["ms"] / mx["steps"]
        out["mixed_fused"] = {"value": mx["rate"], "unit": "Gcell-steps/s", "steps": mx["steps"],
                              "ms_per_step": mms, "speedup_vs_fp64_fused": mx["rate"] / fused["rate"],
                              "algorithmic_bytes_per_cell_step": MIXED_BYTES,
                              "achieved_GBps": MIXED_BYTES * anat.ncell / (mms * 1e-3) / 1e9}
    if "fp64_unfused" in results:
        u = results["fp64_unfused"]
        out["fp64_unfused"] = {"value": u["rate"], "unit": "Gcell-steps/s", "steps": u["steps"],
                               "ms_per_step": u["ms"] / u["steps"],
                               "algorithmic_bytes_per_cell_step": UNFUSED_F64_BYTES}
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle on host cores)")
        out["cpu_baseline"] = cpu_baseline()
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
