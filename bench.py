#!/usr/bin/env python3
"""Throughput of Cardioid's electrics step on one MI355X.

Headline workload: 256x256x512 grid (33.5M points) with a truncated
ellipsoidal-shell anatomy at ~60% tissue (endo/mid/epi by transmural depth),
fibre conductivity with the helix angle rotating -60..+60 deg epi->endo,
8 seeded endocardial stimuli, dt = 0.02 ms, plain FP64 (the drop-in
arithmetic), one fused launch per step.  A "step" is one omp-loop step over
every tissue cell (stimulus, TT06Dev reaction, FGR stencil, range check, Vm
update, activation sensors).

Beside it, on the same anatomy: the mixed-precision fused step and the
reference-structured unfused FP64 step; then 64 batched slabs (mixed, graph
replayed, autotuned launch) and a served request stream (p50/p99 latency).

Prints one JSON line.  value = tissue cell-steps per second of the fused
FP64 engine (Gcell-steps/s), timed with HIP events on the engine's stream
over exactly --steps steps after --warmup untimed ones.
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
# Algorithmic bytes per tissue cell-step of the path (the per-unit figures
# the roofline is quoted on): fused FP64 = 19 FP64 states read+written (304),
# Vm read+written (16), 18 float weights and the double centre weight (80).
FUSED_F64_BYTES = 400
# what the fused FP64 kernel streams: the centre weight is recomputed from the
# float weights, bit for bit, instead of read
FUSED_F64_STREAMED = 392
UNFUSED_F64_BYTES = 480
# mixed: 17 FP32 states r+w (136), Ca_i/Na_i FP64 r+w (32), Vm r+w (16), 19 FP32
# weights (76)
MIXED_BYTES = 260
# what the mixed kernel streams: the compact weights, 18 FP16 differences (36)
# and the tile's 18 float references shared by 64 cells (1.125)
MIXED_STREAMED = 221.125
DT = 0.02


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    """The host CPU as /proc/cpuinfo names it, with family, model, stepping
    and clock: a virtualised host may report only a short model name."""
    info = {}
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if ":" not in line:
                    if info:
                        break            # the first processor's block is enough
                    continue
                k, v = (x.strip() for x in line.split(":", 1))
                info.setdefault(k, v)
    except OSError:
        return "unknown"
    name = info.get("model name", "unknown")
    extra = [f"{k} {info[k]}" for k in ("cpu family", "model", "stepping") if k in info]
    if "cpu MHz" in info:
        extra.append(f"{float(info['cpu MHz']):.0f} MHz")
    if extra:
        name += " (" + ", ".join(extra) + ")"
    return name


def host_threads():
    """The host threads this run may use: OMP_NUM_THREADS where the host
    sets it (a GPU box's share of its CPUs), else every CPU."""
    try:
        n = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        n = 0
    return n if n > 0 else (os.cpu_count() or 1)


def _time_oracle(sim, budget_s, chunk=5):
    sim.run(2)
    steps = 0
    t0 = time.perf_counter()
    while True:
        sim.run(chunk)
        steps += chunk
        el = time.perf_counter() - t0
        if el > budget_s:
            return steps, el


def cpu_baseline(budget_s: float = 12.0):
    """The CPU path on a bounded sample of the same workload (a 96x96x192
    ellipsoidal shell from the same generator, 1.06 M tissue cells, as many
    steps as fit in `budget_s`), extrapolated per cell-step to the 20.1 M-cell
    shell.  Two forms:
      - port: the oracle (the C/OpenMP restatement of the reference CPU
        path) with every loop threaded;
      - reference: the same loop with the reaction run by the reference's own
        kernels (TT06NonGates.c update_nonGate and the TT06GatesSimd.c v2
        gates, built from source into oracle/_ref), (i) as the reference runs
        it, one thread per reaction group and serial host loops, and (ii)
        threaded over cells; stencil and loop are the restatement's, since
        FGRDiffusionOMP.cc and simulationLoop.cc need <mpi.h>, which the image
        lacks.  Present where oracle/_ref was built."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as OL
    from cardioid_amd import workloads as W
    threads = host_threads()
    a, depth = W.ellipsoid_shell(96, 96, 192, seed=0)
    stims = W.endocardial_point_stimuli(a, depth, 8, seed=0)
    host = (f"{cpu_model()}, nproc {os.cpu_count()}, {threads} OpenMP threads used "
            f"(OMP_NUM_THREADS {os.environ.get('OMP_NUM_THREADS', 'unset')})")
    sim = OL.OracleSim(a, W.CELLTYPES_EMI, DT, W.DIFFUSION_SCALE, stims, nthreads=threads)
    steps, el = _time_oracle(sim, budget_s)
    sim.close()
    out = {"value": a.ncell * steps / el / 1e9, "unit": "Gcell-steps/s", "cores": threads,
           "kind": "port", "cpu_model": cpu_model(), "nproc": os.cpu_count(),
           "sample": f"96x96x192 ellipsoidal shell ({a.ncell} tissue cells, the 256x256x512 "
                     f"workload's generator), {steps} steps in {el:.1f} s, oracle/liboracle.so; "
                     f"rate per tissue cell-step extrapolated to the 20.1 M-cell shell; {host}"}
    if OL.ref_available():
        ref = {}
        for form, (rt, pt) in (("i_as_is", (1, 1)), ("ii_threaded", (threads, threads))):
            sim = OL.OracleSim(a, W.CELLTYPES_EMI, DT, W.DIFFUSION_SCALE, stims,
                               nthreads=threads)
            sim.use_reference_reaction(rt, pt)
            steps, el = _time_oracle(sim, budget_s / 2, chunk=1 if rt == 1 else 5)
            sim.close()
            ref[form] = {"value": a.ncell * steps / el / 1e9, "steps": steps, "seconds": el,
                         "reaction_threads": rt}
        ref.update(unit="Gcell-steps/s", kind="reference", cores=threads,
                   sample="the same shell and loop, the reaction by the reference's own "
                          "update_nonGate and v2 gate kernels (oracle/_ref); (i) one thread per "
                          "reaction group and serial integrator and sensors, as the reference "
                          "runs them; (ii) every loop threaded")
        out["reference_kernels"] = ref
    return out


def load_profile(stem):
    path = os.path.join(ROOT, "profiles", stem + "_traffic.json")
    if not os.path.exists(path):
        return {}
    with open(path) as fh:
        return json.load(fh)


def load_traffic(stem):
    return load_profile(stem).get("bytes_per_launch")


def valu(stem, ncell, ms_per_step, peak_tflops=78.6, simds=1024, clock_ghz=2.4):
    """Arithmetic beside the HBM roofline: the path's 1,060 flop per cell-step
    (cardioid_amd/flops.py) against the FP64 vector peak, and the VALU issue fraction from rocprof's
    VALU instructions per wave (4 SIMD cycles each) at this run's step time."""
    from cardioid_amd.flops import flop_per_cell_step
    FLOP_PER_CELL_STEP = flop_per_cell_step()
    p = load_profile(stem)
    rate = ncell / (ms_per_step * 1e-3)
    out = {"flop_per_cell_step": FLOP_PER_CELL_STEP,
           "achieved_tflops": FLOP_PER_CELL_STEP * rate / 1e12, "peak_tflops": peak_tflops,
           "flop_frac": FLOP_PER_CELL_STEP * rate / (peak_tflops * 1e12)}
    ipw = p.get("valu_instructions_per_wave")
    if ipw:
        waves = (ncell + 63) // 64
        out["valu_instructions_per_wave"] = ipw
        out["issue_frac_at_peak_clock"] = (4 * ipw * waves /
                                           (simds * clock_ghz * 1e9 * ms_per_step * 1e-3))
    return out


def stream_ceiling(hip, cells, reps=20):
    """The device's HBM rates for the fused FP64 kernel's state pattern (19
    FP64 fields per cell in 4096-cell runs, the fields in pairs, one 16-byte
    access per lane and pair) over the same number of cells: read only,
    write only, and read and written in place (cdh_stream_rate).  The 8 TB/s peak is not
    reachable for writes; these are the practical ceilings."""
    out = {p + "_GBps": hip.stream_rate(p, cells, reps) for p in ("read", "write", "inplace")}
    out["pattern"] = ("19 FP64 fields per cell in 4096-cell runs with the fields in pairs, one "
                      "non-temporal 16-byte access per lane and pair (8 bytes for the odd field), "
                      f"as the FP64 step streams its states; {cells} cells, {reps} passes each")
    return out


def fused_run(hip, slab, ncell, mode, steps, warmup, tune_steps, graph, sensors=1,
              prebuild=True, exact=False):
    """One slab in the engine: autotuned launch, graph replay, HIP-event time.
    The set-up (autotune, then one graph length of steps that captures the
    graph, plus one step when the warm-up count is odd) runs before the
    warm-up, so that the timed steps start on a graph boundary with the graph
    built: the timed region replays graphs and builds none."""
    cells, box, _ = slab.extent()
    b = hip.Batch(DT, cells, box, slab_capacity=1, mode=mode, sensors=sensors,
                  exact_functions=exact, jhTauSmooth=1 if exact else 0)
    sid = b.add(slab)
    b.sync()
    best, table = b.autotune(tune_steps) if tune_steps > 0 else (None, [])
    b.set_graph(graph)
    setup = (graph + (warmup % 2)) if graph > 0 and prebuild else 0
    b.run(setup)
    b.run(warmup)
    b.sync()
    ms = b.run_timed(steps)
    vm = b.vm(sid)
    act = b.activation(sid)[1]
    t90 = b.t90(sid) if sensors & 2 else None
    rng = b.range_report(sid)
    launch = b.launch()
    b.close()
    return dict(ms=ms, steps=steps, rate=ncell * steps / (ms * 1e-3) / 1e9, vm=vm, act=act,
                t90=t90, launch=launch, tune=table, range=rng, setup_steps=setup)


def launch_json(launch):
    return {"waves": launch[0], "block": launch[1], "occupancy": launch[2],
            "tile_order": launch[3]}


def tune_json(table):
    """the autotuner's table: (workgroup size, occupancy cap, tile order, ms
    per step) per candidate"""
    return [[r[1], r[2], r[3], round(r[4], 5)] for r in table]


def accuracy(np, hip, slab, ncell, steps, graph):
    """Mixed against FP64 on the same anatomy, the same number of steps:
    L-infinity Vm at the end, activation-time error, APD90 error.  The FP64
    run is the fused engine with fused multiply-adds in the fits, within
    1e-14 of the oracle (tests/test_gpu_parity.py); the oracle itself would
    take hours at 20.1 M cells."""
    runs = {}
    for mode in ("fp64_fma", "mixed_fused"):
        t = time.perf_counter()
        runs[mode] = fused_run(hip, slab, ncell, mode, steps, 0, 0, graph, sensors=3,
                               prebuild=False)
        log(f"accuracy: {mode} {steps} steps ({time.perf_counter() - t:.1f} s)")
    f, m = runs["fp64_fma"], runs["mixed_fused"]
    on_f, on_m = f["act"] >= 0, m["act"] >= 0
    both = on_f & on_m
    dact = np.abs(m["act"] - f["act"])[both]
    apd_f, apd_m = f["t90"] - f["act"], m["t90"] - m["act"]
    rep = both & (f["t90"] > 0) & (m["t90"] > 0)
    rel = np.abs(apd_m - apd_f)[rep] / apd_f[rep]
    return {"reference": "fp64_fma (fused FP64, within 1e-14 of the oracle)", "steps": steps,
            "t_ms": steps * DT,
            "linf_vm_mV": float(np.abs(m["vm"] - f["vm"]).max()),
            "activation_max_abs_ms": float(dact.max()) if dact.size else None,
            "activation_identical_frac": float(np.mean(dact == 0)) if dact.size else None,
            "activated_fp64": int(on_f.sum()),
            "activated_in_one_run_only": int((on_f != on_m).sum()),
            "apd90_rel_max": float(rel.max()) if rel.size else None,
            "apd90_rel_mean": float(rel.mean()) if rel.size else None,
            "repolarised_fp64": int((f["t90"] > 0).sum()),
            "repolarised_in_one_run_only": int(((f["t90"] > 0) != (m["t90"] > 0)).sum())}


def batched(hip, W, steps, tune_steps, graph):
    cases = W.batched_slabs()
    slabs = [hip.Slab(c.anatomy, ["endoCellML"], W.DIFFUSION_SCALE, c.stimuli, scales=c.scales)
             for c in cases]
    ext = [s.extent() for s in slabs]
    ncell = sum(s.ncell for s in slabs)
    b = hip.Batch(DT, sum(e[0] for e in ext), sum(e[1] for e in ext), slab_capacity=len(slabs),
                  mode="mixed_fused")
    t = time.perf_counter()
    ids = [b.add(s) for s in slabs]
    b.sync()
    log(f"batched: {len(slabs)} slabs, {ncell} cells, added in {time.perf_counter() - t:.1f} s")
    best, table = b.autotune(tune_steps)
    b.set_graph(graph)
    b.run(graph)
    b.sync()
    ms = b.run_timed(steps)
    act = [b.activation(i)[1] for i in ids]
    launch = b.launch()
    b.close()
    return {"value": ncell * steps / (ms * 1e-3) / 1e9, "unit": "Gcell-steps/s",
            "slabs": len(slabs), "tissue_cells": ncell, "steps": steps,
            "ms_per_step": ms / steps,
            "launch": launch_json(launch), "autotune": tune_json(table),
            "graph_steps": graph,
            "slabs_activated": int(sum((a >= 0).mean() > 0.5 for a in act)),
            "achieved_GBps": MIXED_BYTES * ncell * steps / (ms * 1e-3) / 1e9,
            "streamed_GBps": MIXED_STREAMED * ncell * steps / (ms * 1e-3) / 1e9}


def shards(hip, W, anat, stims, parts, mode, launch, steps, graph):
    """The whole anatomy cut into `parts` x-slabs, as `parts` GPUs would hold
    it: all shares in one engine with their halo planes exchanged after every
    step (the device-local form of the exchange), and the largest share alone
    (one GPU's work per step in a `parts`-GPU run, halo aside)."""
    shares = W.split_x(anat, parts)
    slabs = [hip.Slab(s.anatomy, W.CELLTYPES_EMI, W.DIFFUSION_SCALE, stims,
                      remote_gid=s.remote_gid, remote_sigma=s.remote_sigma) for s in shares]
    ext = [s.extent() for s in slabs]
    b = hip.Batch(DT, sum(e[0] for e in ext), sum(e[1] for e in ext), slab_capacity=len(slabs),
                  mode=mode)
    ids = [b.add(s) for s in slabs]
    halo = 0
    for k in range(1, len(ids)):
        halo += b.link_halo(ids[k], ids[k - 1]) + b.link_halo(ids[k - 1], ids[k])
    b.set_launch(*launch)
    b.set_graph(graph)
    b.run(20)
    b.sync()
    ms_all = b.run_timed(steps) / steps
    b.close()
    big = max(range(len(slabs)), key=lambda k: slabs[k].ncell)
    e = ext[big]
    b = hip.Batch(DT, e[0], e[1], slab_capacity=1, mode=mode)
    b.add(slabs[big])
    b.set_launch(*launch)
    b.set_graph(graph)
    b.run(20)
    b.sync()
    ms_one = b.run_timed(steps) / steps
    b.close()
    return {"parts": len(slabs), "mode": mode, "steps": steps,
            "ms_per_step_all_shares_with_halo_exchange": ms_all,
            "halo_points_exchanged_per_step": halo,
            "largest_share_cells": slabs[big].ncell,
            "largest_share_x_planes": list(shares[big].x_range),
            "ms_per_step_largest_share_alone": ms_one,
            "largest_share_Gcell_steps_s": slabs[big].ncell / (ms_one * 1e-3) / 1e9,
            "largest_share_halo_bytes_per_step": 8 * slabs[big].nremote}


def serving(hip, W, nreq, rate, hbm_gb, graph, max_active_cells=None):
    from cardioid_amd import serve
    reqs = W.request_stream(nreq, rate_hz=rate, seed=4)
    cells, box = serve.arena_for_budget(int(hbm_gb * 2 ** 30))
    b = hip.Batch(DT, cells, box, slab_capacity=1024, mode="mixed_fused", sensors=3)
    b.set_launch(3, 64, 0, -1)
    b.set_graph(graph)
    st = serve.Server(b, chunk=100, max_active_cells=max_active_cells).serve(reqs)
    b.close()
    s = st.summary()
    s.update({"rate_hz": rate, "hbm_budget_gb": hbm_gb, "arena_cells": cells,
              "max_active_cells": max_active_cells,
              "request_dims": "each axis log-uniform 32..128, duration 200-1000 ms, "
                              "scales U[0,1], mixed precision"})
    return s


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--grid", type=str, default="256,256,512")
    ap.add_argument("--unfused-steps", type=int, default=500)
    ap.add_argument("--exact-unfused-steps", type=int, default=300,
                    help="the unfused loop in configuration A (TT06Dev tolerance 0, the "
                         "factory's default deck) for this many steps")
    ap.add_argument("--exact-fused-steps", type=int, default=1000,
                    help="the fused FP64 engine in configuration A for this many steps")
    ap.add_argument("--mixed-steps", type=int, default=5000)
    ap.add_argument("--exact-steps", type=int, default=2000)
    ap.add_argument("--tune-steps", type=int, default=40)
    ap.add_argument("--graph", type=int, default=10)
    ap.add_argument("--batched-steps", type=int, default=10000)
    ap.add_argument("--serve-requests", type=int, default=40)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-hbm-gb", type=float, default=24.0)
    ap.add_argument("--serve-cap", type=int, default=6_000_000,
                    help="also serve the stream with at most this many cells in flight")
    ap.add_argument("--steady-requests", type=int, default=100,
                    help="a second stream below capacity: this many requests ...")
    ap.add_argument("--steady-rate", type=float, default=1.5, help="... at this rate (req/s)")
    ap.add_argument("--accuracy-steps", type=int, default=25000,
                    help="mixed against FP64 on the headline anatomy over this many steps")
    ap.add_argument("--shard-parts", type=int, default=8)
    ap.add_argument("--shard-steps", type=int, default=1000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: every box in this setup has a single GPU and the "
              "path is not sharded across GPUs here (--gpus must be 1).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from cardioid_amd import hip, workloads as W
    nx, ny, nz = (int(v) for v in args.grid.split(","))
    log(f"generating {nx}x{ny}x{nz} ellipsoidal shell")
    t = time.perf_counter()
    anat, depth = W.ellipsoid_shell(nx, ny, nz, seed=0)
    stims = W.endocardial_point_stimuli(anat, depth, 8, seed=0)
    log(f"{anat.ncell} tissue cells ({anat.ncell / (nx * ny * nz):.3f} of the grid) "
        f"in {time.perf_counter() - t:.1f} s")
    slab = hip.Slab(anat, W.CELLTYPES_EMI, W.DIFFUSION_SCALE, stims)

    res = {}
    for mode, steps in (("fp64_fma", args.steps), ("mixed_fused", args.mixed_steps),
                        ("fp64_fused", args.exact_steps)):
        if steps <= 0:
            continue
        t = time.perf_counter()
        r = fused_run(hip, slab, anat.ncell, mode, steps, args.warmup, args.tune_steps, args.graph)
        res[mode] = r
        log(f"{mode}: launch {r['launch']}, {steps} steps in {r['ms']:.1f} ms -> "
            f"{r['rate']:.3f} Gcell-steps/s ({time.perf_counter() - t:.1f} s incl. set-up)")
    if args.exact_fused_steps > 0:
        t = time.perf_counter()
        r = fused_run(hip, slab, anat.ncell, "fp64_fused", args.exact_fused_steps, args.warmup,
                      args.tune_steps, args.graph, exact=True)
        res["fp64_fused_exact"] = r
        log(f"fp64_fused, configuration A: launch {r['launch']}, {r['rate']:.3f} Gcell-steps/s "
            f"({time.perf_counter() - t:.1f} s incl. set-up)")
    if args.unfused_steps > 0:
        sim = hip.Simulation(anat, W.CELLTYPES_EMI, DT, W.DIFFUSION_SCALE, stims,
                             mode="fp64_unfused")
        sim.run(args.warmup)
        sim.sync()
        ms = sim.run_timed(args.unfused_steps)
        sim.close()
        res["fp64_unfused"] = dict(ms=ms, steps=args.unfused_steps,
                                   rate=anat.ncell * args.unfused_steps / (ms * 1e-3) / 1e9)
        log(f"fp64_unfused: {res['fp64_unfused']['rate']:.3f} Gcell-steps/s")
    if args.exact_unfused_steps > 0:
        sim = hip.Simulation(anat, W.CELLTYPES_EMI, DT, W.DIFFUSION_SCALE, stims,
                             mode="fp64_unfused", exact_functions=True, jhTauSmooth=1)
        sim.run(args.warmup)
        sim.sync()
        ms = sim.run_timed(args.exact_unfused_steps)
        sim.close()
        res["fp64_unfused_exact"] = dict(
            ms=ms, steps=args.exact_unfused_steps,
            rate=anat.ncell * args.exact_unfused_steps / (ms * 1e-3) / 1e9)
        log(f"fp64_unfused, configuration A: {res['fp64_unfused_exact']['rate']:.3f} "
            f"Gcell-steps/s")

    f64, mx = res["fp64_fma"], res["mixed_fused"]
    ms_per_step = f64["ms"] / f64["steps"]
    achieved = FUSED_F64_BYTES * anat.ncell / (ms_per_step * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": f64["rate"],
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
                               "fused step (fit polynomials with FMA; Vm parity with the "
                               "reference path within 1e-9)",
                   "grid_points": nx * ny * nz, "tissue_cells": anat.ncell, "dt_ms": DT,
                   "mode": "fp64_fma",
                   "launch": launch_json(f64["launch"]),
                   "graph_steps": args.graph,
                   "setup_steps_before_warmup": f64["setup_steps"]},
        "autotune": {"fp64_fma": tune_json(f64["tune"]), "mixed_fused": tune_json(mx["tune"]),
                     "columns": ["block", "occupancy cap", "tile order", "ms per step"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": load_traffic("engine_f64_fma_shell256x256x512"),
                     "kernel": "k_engine_f64", "algorithmic_bytes_per_cell_step": FUSED_F64_BYTES,
                     "streamed_bytes_per_cell_step": FUSED_F64_STREAMED,
                     "streamed_frac": achieved * FUSED_F64_STREAMED / FUSED_F64_BYTES
                     / HBM_PEAK_GBS},
        "valu": valu("engine_f64_fma_shell256x256x512", anat.ncell, ms_per_step),
    }
    ceil = stream_ceiling(hip, anat.ncell)
    out["stream_ceiling"] = ceil
    streamed = achieved * FUSED_F64_STREAMED / FUSED_F64_BYTES
    out["roofline"]["streamed_GBps"] = streamed
    out["roofline"]["streamed_vs_inplace_stream"] = streamed / ceil["inplace_GBps"]
    log(f"stream ceiling: read {ceil['read_GBps']:.0f}, write {ceil['write_GBps']:.0f}, in place "
        f"{ceil['inplace_GBps']:.0f} GB/s; the FP64 step streams {streamed:.0f} GB/s")
    mms = mx["ms"] / mx["steps"]
    both = (f64["act"] >= 0) & (mx["act"] >= 0)
    dropin = res["fp64_unfused"]["rate"] if "fp64_unfused" in res else None
    refr = res["fp64_fused"]["rate"] if "fp64_fused" in res else None
    out["mixed_fused"] = {
        "value": mx["rate"], "unit": "Gcell-steps/s", "steps": mx["steps"], "ms_per_step": mms,
        # one definition: mixed cell-steps/s over the FP64 path's, same anatomy
        # and step; the fused FP64 step is the headline ratio
        "speedup_vs_fp64_fused_fma": mx["rate"] / f64["rate"],
        "speedup_vs_fp64_fused_reference_rounding": mx["rate"] / refr if refr else None,
        "speedup_vs_fp64_dropin": mx["rate"] / dropin if dropin else None,
        "launch": launch_json(mx["launch"]),
        "algorithmic_bytes_per_cell_step": MIXED_BYTES,
        "achieved_GBps": MIXED_BYTES * anat.ncell / (mms * 1e-3) / 1e9,
        "streamed_bytes_per_cell_step": MIXED_STREAMED,
        "streamed_GBps": MIXED_STREAMED * anat.ncell / (mms * 1e-3) / 1e9,
        "traffic": load_traffic("engine_mixed_shell256x256x512"),
        "valu": valu("engine_mixed_shell256x256x512", anat.ncell, mms)}
    if args.accuracy_steps > 0:
        acc = accuracy(np, hip, slab, anat.ncell, args.accuracy_steps, args.graph)
        out["mixed_fused"].update(
            linf_vm_mV=acc["linf_vm_mV"], activation_max_abs_ms=acc["activation_max_abs_ms"],
            apd90_rel_max=acc["apd90_rel_max"], error_vs_fp64=acc)
        log(f"accuracy over {acc['t_ms']:.0f} ms: L-inf Vm {acc['linf_vm_mV']:.3e} mV, "
            f"activation {acc['activation_max_abs_ms']} ms, APD90 rel {acc['apd90_rel_max']}")
    out["range_check"] = {"fp64_out_of_range": (f64["range"] or {}).get("count", 0),
                          "mixed_out_of_range": (mx["range"] or {}).get("count", 0)}
    if "fp64_fused" in res:
        ex = res["fp64_fused"]
        out["fp64_reference_rounding"] = {
            "value": ex["rate"], "unit": "Gcell-steps/s", "steps": ex["steps"],
            "ms_per_step": ex["ms"] / ex["steps"],
            "launch": launch_json(ex["launch"]),
            "note": "no FMA anywhere: bit for bit the unfused step's rounding",
            "mixed_speedup": mx["rate"] / ex["rate"]}
    if "fp64_unfused" in res:
        u = res["fp64_unfused"]
        out["fp64_unfused"] = {"value": u["rate"], "unit": "Gcell-steps/s", "steps": u["steps"],
                               "ms_per_step": u["ms"] / u["steps"],
                               "algorithmic_bytes_per_cell_step": UNFUSED_F64_BYTES}
    if "fp64_unfused_exact" in res:
        u = res["fp64_unfused_exact"]
        out["fp64_unfused_configuration_A"] = {
            "value": u["rate"], "unit": "Gcell-steps/s", "steps": u["steps"],
            "ms_per_step": u["ms"] / u["steps"],
            "note": "TT06Dev tolerance 0, mod 1 (the reaction factory's default deck): exact "
                    "functions with the device's exp/log/pow, generic gate update; the same "
                    "loop and diffusion as fp64_unfused"}
    if "fp64_fused_exact" in res:
        ex = res["fp64_fused_exact"]
        out["fp64_fused_configuration_A"] = {
            "value": ex["rate"], "unit": "Gcell-steps/s", "steps": ex["steps"],
            "ms_per_step": ex["ms"] / ex["steps"], "launch": launch_json(ex["launch"]),
            "note": "TT06Dev tolerance 0, mod 1 in the fused engine: bit for bit the unfused "
                    "configuration A step"}
        if "fp64_unfused_exact" in res:
            out["fp64_fused_configuration_A"]["speedup_over_unfused"] = (
                ex["rate"] / res["fp64_unfused_exact"]["rate"])
    if args.shard_steps > 0 and args.shard_parts > 1:
        sh = shards(hip, W, anat, stims, args.shard_parts, "fp64_fma", f64["launch"],
                    args.shard_steps, args.graph)
        sh["ms_per_step_whole"] = ms_per_step
        out["shards"] = sh
        log(f"shards: {sh['parts']} x-slabs with halo exchange "
            f"{sh['ms_per_step_all_shares_with_halo_exchange']:.3f} ms/step (whole "
            f"{ms_per_step:.3f}); largest share alone {sh['ms_per_step_largest_share_alone']:.3f} "
            f"ms/step")
    if args.batched_steps > 0:
        out["batched"] = batched(hip, W, args.batched_steps, args.tune_steps, args.graph)
        log(f"batched: {out['batched']['value']:.3f} Gcell-steps/s")
    if args.serve_requests > 0:
        out["serving"] = serving(hip, W, args.serve_requests, args.serve_rate, args.serve_hbm_gb,
                                 args.graph)
        log(f"serving: p50 {out['serving']['p50_latency_s']:.2f} s, "
            f"p99 {out['serving']['p99_latency_s']:.2f} s")
        if args.serve_cap > 0:
            # the same stream with at most serve_cap cells in flight: requests
            # finish in arrival order (lower median, same tail)
            out["serving_capped"] = serving(hip, W, args.serve_requests, args.serve_rate,
                                            args.serve_hbm_gb, args.graph, args.serve_cap)
            log(f"serving, {args.serve_cap} cells in flight at most: p50 "
                f"{out['serving_capped']['p50_latency_s']:.2f} s, p99 "
                f"{out['serving_capped']['p99_latency_s']:.2f} s")
    if args.steady_requests > 0:
        # the same request mix below the GPU's capacity (about 1.8 req/s of
        # this mix at ~19 Gcell-steps/s): latency is admission and shared
        # steps, not one overloaded batch's makespan
        out["serving_steady"] = serving(hip, W, args.steady_requests, args.steady_rate,
                                        args.serve_hbm_gb, args.graph)
        log(f"serving at {args.steady_rate} req/s: p50 "
            f"{out['serving_steady']['p50_latency_s']:.2f} s, p99 "
            f"{out['serving_steady']['p99_latency_s']:.2f} s")
    if not args.no_cpu_baseline:
        log("cpu baselines (oracle, and the reference's reaction kernels, on host cores)")
        out["cpu_baseline"] = cpu_baseline()
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
