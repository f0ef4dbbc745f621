"""Writes tests/golden/niederer_dx0.2_dt0.01.json: the oracle's activation
times (first Vm > 0, ActivationTimeSensor) on the Niederer benchmark slab at
dx = 0.2 mm, dt = 0.01 ms, 60 ms (6000 steps): the nine benchmark points,
the number of activated cells and the sum of all activation times.

    python tests/golden/make_niederer.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle_lib as OL  # noqa: E402
from cardioid_amd import workloads as W  # noqa: E402


def run(dx=0.2, dt=0.01, t_end=60.0, nthreads=0):
    a, stim, idx = W.niederer(dx)
    sim = OL.OracleSim(a, ["epiCellML"], dt, W.DIFFUSION_SCALE, stim, nthreads=nthreads)
    sim.run(int(round(t_end / dt)))
    act = sim.activation()[0]
    sim.close()
    return act, idx


def main():
    act, idx = run()
    out = {"dx": 0.2, "dt": 0.01, "steps": 6000, "points": W.NIEDERER_POINTS,
           "tact0_points": [float(act[i]) for i in idx],
           "activated": int((act >= 0).sum()), "tact0_sum": float(act[act >= 0].sum())}
    with open(os.path.join(HERE, "niederer_dx0.2_dt0.01.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
