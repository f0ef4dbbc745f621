"""Mixed precision against FP64 on the headline anatomy (256x256x512
ellipsoidal shell, 8 endocardial stimuli) over 500 ms (25,000 steps of
0.02 ms): L-infinity Vm error at the end, activation-time error per cell
and APD90 error per cell (APD90 = t90 - activation time, t90 the first time
after activation that Vm falls below Vrest + 0.1 (Vmax - Vrest)).

    python -m cardioid_amd.accuracy [--steps N] [--grid X,Y,Z] [--out FILE]

Prints one JSON line (and writes it to --out).
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import hip, workloads as W

DT = 0.02


def run(slab, mode, steps, chunk=5000):
    cells, box, _ = slab.extent()
    b = hip.Batch(DT, cells, box, slab_capacity=1, mode=mode, sensors=3)
    sid = b.add(slab)
    b.set_launch(*((3, 64) if mode == "mixed_fused" else (2, 64)))
    b.set_graph(10)
    done, ms = 0, 0.0
    while done < steps:
        k = min(chunk, steps - done)
        ms += b.run_timed(k)
        done += k
        print(f"[accuracy] {mode}: {done}/{steps} steps, {ms / done:.3f} ms/step", file=sys.stderr,
              flush=True)
    out = dict(vm=b.vm(sid), act=b.activation(sid)[1], t90=b.t90(sid), ms=ms)
    b.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=25000)
    ap.add_argument("--grid", default="256,256,512")
    ap.add_argument("--out", default="")
    a = ap.parse_args(argv)
    nx, ny, nz = (int(v) for v in a.grid.split(","))
    t = time.perf_counter()
    anat, depth = W.ellipsoid_shell(nx, ny, nz, seed=0)
    stims = W.endocardial_point_stimuli(anat, depth, 8, seed=0)
    slab = hip.Slab(anat, W.CELLTYPES_EMI, W.DIFFUSION_SCALE, stims)
    print(f"[accuracy] {anat.ncell} cells in {time.perf_counter() - t:.1f} s", file=sys.stderr,
          flush=True)
    ref = run(slab, "fp64_fused", a.steps)
    mix = run(slab, "mixed_fused", a.steps)
    act_ok = (ref["act"] >= 0) & (mix["act"] >= 0)
    dact = np.abs(mix["act"] - ref["act"])[act_ok]
    rep = (ref["t90"] > 0) & (mix["t90"] > 0) & act_ok
    apd_r = (ref["t90"] - ref["act"])[rep]
    apd_m = (mix["t90"] - mix["act"])[rep]
    rel = np.abs(apd_m - apd_r) / apd_r
    out = {
        "workload": f"ellipsoidal shell {nx}x{ny}x{nz}, 8 endocardial stimuli, dt {DT} ms",
        "tissue_cells": anat.ncell, "steps": a.steps, "t_end_ms": a.steps * DT,
        "linf_vm_mV": float(np.abs(mix["vm"] - ref["vm"]).max()),
        "activated_fp64": float((ref["act"] >= 0).mean()),
        "activated_same_cells": bool(np.array_equal(ref["act"] >= 0, mix["act"] >= 0)),
        "activation_max_abs_ms": float(dact.max()), "activation_max_abs_dt": float(dact.max() / DT),
        "activation_identical_frac": float((dact == 0).mean()),
        "repolarised_fp64": float((ref["t90"] > 0).mean()),
        "repolarised_same_cells": bool(np.array_equal(ref["t90"] > 0, mix["t90"] > 0)),
        "apd90_fp64_ms": [float(apd_r.min()), float(apd_r.max())],
        "apd90_rel_err_max": float(rel.max()), "apd90_rel_err_mean": float(rel.mean()),
        "ms_per_step": {"fp64_fused": ref["ms"] / a.steps, "mixed_fused": mix["ms"] / a.steps},
    }
    line = json.dumps(out)
    print(line, flush=True)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(line + "\n")


if __name__ == "__main__":
    main()
