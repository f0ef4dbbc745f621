"""Profiling driver: the fused engine on the 256x256x512 ellipsoidal-shell
workload for a few steps, preceded by a plain 512 MiB device-to-device copy
whose known byte count calibrates rocprofv3's FETCH_SIZE / WRITE_SIZE.

    python -m cardioid_amd.profile_step [variant] [steps] [mode]

Meant to run under rocprofv3 (kernel trace, or one --pmc counter per run).
"""
from __future__ import annotations

import sys


def main(argv):
    import torch

    from cardioid_amd import hip, workloads as W

    variant = int(argv[0]) if len(argv) > 0 else 0
    steps = int(argv[1]) if len(argv) > 1 else 20
    mode = argv[2] if len(argv) > 2 else "fp64_fused"
    x = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    anat, depth = W.ellipsoid_shell(256, 256, 512, seed=0)
    stims = W.endocardial_point_stimuli(anat, depth, 8, seed=0)
    sim = hip.Simulation(anat, W.CELLTYPES_EMI, 0.02, W.DIFFUSION_SCALE, stims, mode=mode,
                         variant=variant)
    sim.run(steps)
    sim.sync()
    print(f"profile_step: {mode} variant {variant}, {steps} steps on {anat.ncell} cells",
          flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
