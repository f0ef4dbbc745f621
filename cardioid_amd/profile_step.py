"""Profiling driver: the fused engine on the 256x256x512 ellipsoidal-shell
workload (or the 64 batched slabs) for a few steps, preceded by a plain
512 MiB device-to-device copy whose known byte count calibrates rocprofv3's
FETCH_SIZE / WRITE_SIZE.

    python -m cardioid_amd.profile_step MODE STEPS WAVES BLOCK OCC [shell|batch]

MODE is fp64_fused, fp64_fma or mixed_fused (the fused engine; WAVES/BLOCK/OCC
the launch configuration: register target, workgroup size, occupancy cap), or
fp64_unfused (the reference loop over the drop-in objects: reaction, scatter,
FGR stencil, integrator and sensor kernels; WAVES/BLOCK/OCC ignored), or
fp64_unfused_exact (the same loop with the reaction factory's default deck,
TT06Dev tolerance 0: the exact functions), or fp64_fused_exact (that deck in
the fused FP64 engine).  Meant
to run under rocprofv3 (kernel trace, or one --pmc counter per run).
"""
from __future__ import annotations

import sys


def main(argv):
    import torch

    from cardioid_amd import hip, workloads as W

    mode = argv[0] if len(argv) > 0 else "fp64_fused"
    steps = int(argv[1]) if len(argv) > 1 else 20
    waves = int(argv[2]) if len(argv) > 2 else 3
    block = int(argv[3]) if len(argv) > 3 else 64
    occ = int(argv[4]) if len(argv) > 4 else 0
    work = argv[5] if len(argv) > 5 else "shell"
    x = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    if mode in ("fp64_unfused", "fp64_unfused_exact"):
        anat, depth = W.ellipsoid_shell(256, 256, 512, seed=0)
        stims = W.endocardial_point_stimuli(anat, depth, 8, seed=0)
        kw = dict(exact_functions=True, jhTauSmooth=1) if mode.endswith("exact") else {}
        sim = hip.Simulation(anat, W.CELLTYPES_EMI, 0.02, W.DIFFUSION_SCALE, stims,
                             mode="fp64_unfused", **kw)
        sim.run(steps)
        sim.sync()
        sim.close()
        print(f"profile_step: {mode} shell, {steps} steps on {anat.ncell} cells", flush=True)
        return
    if work == "shell":
        anat, depth = W.ellipsoid_shell(256, 256, 512, seed=0)
        stims = W.endocardial_point_stimuli(anat, depth, 8, seed=0)
        slabs = [hip.Slab(anat, W.CELLTYPES_EMI, W.DIFFUSION_SCALE, stims)]
    else:
        slabs = [hip.Slab(c.anatomy, ["endoCellML"], W.DIFFUSION_SCALE, c.stimuli, scales=c.scales)
                 for c in W.batched_slabs()]
    ext = [s.extent() for s in slabs]
    exact = mode == "fp64_fused_exact"
    b = hip.Batch(0.02, sum(e[0] for e in ext), sum(e[1] for e in ext), slab_capacity=len(slabs),
                  mode="fp64_fused" if exact else mode, exact_functions=exact,
                  jhTauSmooth=1 if exact else 0)
    for s in slabs:
        b.add(s)
    b.set_launch(waves, block, occ)
    b.run(steps)
    b.sync()
    ncell = sum(s.ncell for s in slabs)
    print(f"profile_step: {mode} {work} launch ({waves}, {block}, {occ}), {steps} steps on "
          f"{ncell} cells",
          flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
