"""Times the reference's CPU path on this host (test infrastructure: the
CPU baseline, not the product).

The omp-loop step (elec/simulationLoop.cc:170-241) runs in the oracle's loop
with the reaction done by the reference's own kernels, update_nonGate
(elec/TT06NonGates.c:374-563) and the v2 gate kernels (elec/TT06GatesSimd.c)
through TT06Func::updateGateFast, compiled from source into
oracle/_ref/libcardioid_ref.so (make -C oracle ref).  The rest of the step
(scatter, FGR stencil, integrator, sensors) is the oracle's restatement:
FGRDiffusionOMP.cc and simulationLoop.cc include <mpi.h>, which this image
lacks, so they cannot be compiled here.  The loop with the reference's
reaction equals the oracle's own loop bit for bit
(tests/test_oracle_reaction.py::test_loop_with_the_reference_reaction_equals_the_oracle).

Two timings, as the reference runs and as it would run threaded:
  (i)  as is: the reaction on one thread per reaction group (TT06Dev_Reaction::calc,
       TT06Dev_Reaction.cc:393-407), iStim fill, integrator and sensors on
       one thread (HOST_PARALLEL_FORALL / DEVICE_PARALLEL_FORALL on the host,
       elec/DeviceFor.hh), scatter and stencil OpenMP-parallel
       (FGRDiffusionOMP.cc:75-131);
  (ii) the same kernels with every loop OpenMP-parallel over cells (the
       reaction split in 12-cell multiples, where the v2 loops' per-group
       rules fall as in the whole group).
Also the oracle's own kernels (the port) in the same loop, threaded.

    python oracle/time_reference.py [--threads N] [--cfg0-steps K] [--cfg1-steps K]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def time_run(sim, steps: int) -> float:
    sim.run(1)                  # first touch of every buffer outside the timing
    t = time.perf_counter()
    sim.run(steps)
    return time.perf_counter() - t


def measure(anat, celltypes, stims, threads: int, steps: dict, label: str, log=None):
    import oracle_lib as OL
    from cardioid_amd import workloads as W
    log = log or (lambda m: print(m, file=sys.stderr, flush=True))
    out = {}
    for form in ("reference_i", "reference_ii", "port"):
        if form.startswith("reference") and not OL.ref_available():
            continue
        n = steps[form]
        sim = OL.OracleSim(anat, celltypes, 0.02, W.DIFFUSION_SCALE, stims, nthreads=threads)
        if form == "reference_i":
            sim.use_reference_reaction(1, 1)
        elif form == "reference_ii":
            sim.use_reference_reaction(threads, threads)
        dt = time_run(sim, n)
        sim.close()
        rate = anat.ncell * n / dt
        out[form] = {"cell_steps_per_s": rate, "steps_timed": n, "seconds": dt}
        log(f"[time_reference] {label} {form}: {n} steps of {anat.ncell} cells in {dt:.2f} s "
            f"-> {rate / 1e6:.2f} Mcell-steps/s")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description="time the reference's CPU path")
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--cfg0-steps", type=int, default=100)
    ap.add_argument("--cfg1-steps", type=int, default=20)
    ap.add_argument("--skip-cfg1", action="store_true")
    a = ap.parse_args(argv)
    from cardioid_amd import workloads as W
    res = {"host": {"cpu_model": cpu_model(), "nproc": os.cpu_count(), "threads": a.threads}}
    b = W.brick(64, 64, 64)
    res["cfg0"] = measure(b, ["endoCellML"], [W.box_stimulus(xMax=4, nx=64, ny=64, nz=64)],
                          a.threads, {"reference_i": a.cfg0_steps, "reference_ii": a.cfg0_steps,
                                      "port": a.cfg0_steps}, "cfg0 64^3")
    res["cfg0"]["note"] = (f"64^3 all-tissue endocardial brick; {a.cfg0_steps} of the 1000 steps "
                           "timed after one untimed step, rates per tissue cell-step")
    if not a.skip_cfg1:
        shell, depth = W.ellipsoid_shell(256, 256, 512, seed=0)
        st = W.endocardial_point_stimuli(shell, depth, 8, seed=0)
        res["cfg1"] = measure(shell, W.CELLTYPES_EMI, st, a.threads,
                              {"reference_i": a.cfg1_steps, "reference_ii": a.cfg1_steps,
                               "port": a.cfg1_steps}, "cfg1 256x256x512")
        res["cfg1"]["note"] = (f"256x256x512 ellipsoidal shell, {shell.ncell} tissue cells; "
                               f"{a.cfg1_steps} of the 5000 steps timed, the rate extrapolated")
    print(json.dumps(res, indent=1))
    return res


if __name__ == "__main__":
    main()
