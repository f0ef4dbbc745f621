"""Mixed-precision fused engine against the FP64 oracle on the same inputs.

The mixed form keeps Vm, Ca_i and Na_i in FP64 and the other states and
the stencil in FP32.  Bar, against the oracle's FP64 run over 500 ms:
activation times within one time step, APD90 within 1 %, the same cells
activated and repolarised, and the L-infinity Vm error reported and
bounded.  On the full 256x256x512 anatomy, where the oracle takes hours,
the mixed run is held to the FP64 engine (itself within 1e-14 of the
oracle) over 5000 steps.
"""
import sys

import numpy as np
import pytest

from cardioid_amd import hip, workloads as W

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
DT = 0.02


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    yield


def run(a, stims, mode, n, scales=None):
    sim = hip.Simulation(a, W.CELLTYPES_EMI, DT, W.DIFFUSION_SCALE, stims, mode=mode, sensors=3,
                         scales=scales)
    sim.run(n)
    out = dict(vm=sim.vm(), act=sim.activation()[1], t90=sim.t90(), state=sim.state())
    sim.close()
    return out


def errors_vs(ref, mix):
    """(activation |dt| max, fraction identical, APD90 rel max, rel mean,
    repolarised fraction, L-inf Vm)"""
    dact = np.abs(mix["act"] - ref["act"])
    ok = (ref["t90"] > 0) & (mix["t90"] > 0)
    apd_ref = ref["t90"] - ref["act"]
    apd_mix = mix["t90"] - mix["act"]
    rel = np.abs(apd_mix - apd_ref)[ok] / apd_ref[ok] if ok.any() else np.zeros(1)
    return (dact.max(), np.mean(dact == 0), rel.max(), rel.mean(), ok.mean(),
            np.abs(mix["vm"] - ref["vm"]).max())


def test_mixed_against_the_fp64_oracle_over_500_ms():
    """cfg2's measurement on the oracle's scale: the 32x28x40 ellipsoidal
    shell (endo/mid/epi, rotating fibres, 8 endocardial stimuli), 25000
    steps of 0.02 ms, mixed engine against the oracle's FP64 run."""
    import oracle_lib as OL
    nsteps = 25000
    say(f"mixed vs FP64 oracle: shell 32x28x40, 8 endocardial stimuli, {nsteps} steps")
    a, depth = W.ellipsoid_shell(32, 28, 40, seed=0)
    stims = W.endocardial_point_stimuli(a, depth, 8, seed=0)
    mix = run(a, stims, "mixed_fused", nsteps)
    ora = OL.OracleSim(a, W.CELLTYPES_EMI, DT, W.DIFFUSION_SCALE, stims)
    for k in range(5):
        ora.run(nsteps // 5)
        say(f"  oracle {(k + 1) * nsteps // 5} steps")
    ref = dict(vm=ora.vm(), act=ora.activation()[1], t90=ora.t90())
    ora.close()
    dact, same, apd_max, apd_mean, repol, linf = errors_vs(ref, mix)
    say(f"  activation |dt| max {dact:.4f} ms (dt {DT}), identical {same:.4f}")
    say(f"  APD90 rel err max {apd_max:.2e} mean {apd_mean:.2e}, repolarised {repol:.3f}")
    say(f"  L-inf Vm after {nsteps * DT:.0f} ms: {linf:.3e} mV")
    assert (ref["act"] >= 0).all() and (mix["act"] >= 0).all()
    assert dact <= DT * 1.0000001
    assert repol > 0.99 and np.array_equal(ref["t90"] > 0, mix["t90"] > 0)
    assert apd_max < 0.01
    assert linf < 1.0


def test_mixed_on_the_full_anatomy_against_fp64():
    """cfg2's anatomy at its own size: 256x256x512 ellipsoidal shell (20.1 M
    tissue cells), 5000 steps (100 ms) of the mixed engine against the FP64
    engine with fused multiply-adds in the fits (within 1e-14 of the oracle):
    the same cells activated, activation times within one step, Vm within a
    bound while the tissue is in its plateau."""
    nsteps = 5000
    say("mixed vs fp64_fma on 256x256x512 (20.1 M cells), 5000 steps")
    a, depth = W.ellipsoid_shell(256, 256, 512, seed=0)
    stims = W.endocardial_point_stimuli(a, depth, 8, seed=0)
    out = {}
    for mode in ("fp64_fma", "mixed_fused"):
        sim = hip.Simulation(a, W.CELLTYPES_EMI, DT, W.DIFFUSION_SCALE, stims, mode=mode)
        for k in range(5):
            sim.run(nsteps // 5)
            sim.sync()
            say(f"  {mode} {(k + 1) * nsteps // 5} steps")
        out[mode] = dict(vm=sim.vm(), act=sim.activation()[1])
        sim.close()
    ref, mix = out["fp64_fma"], out["mixed_fused"]
    end = nsteps * DT
    on, on_mix = ref["act"] >= 0, mix["act"] >= 0
    both = on & on_mix
    dact = np.abs(mix["act"] - ref["act"])[both]
    # a cell activated in one run only was reached in the run's last step
    only = on ^ on_mix
    last = np.where(on, ref["act"], mix["act"])[only]
    dv = np.abs(mix["vm"] - ref["vm"])
    plateau = both & (ref["act"] < end - 10.0)
    say(f"  activated {on.mean():.4f}; activation |dt| max {dact.max():.4f} ms, identical "
        f"{np.mean(dact == 0):.4f}; in one run only {only.sum()}")
    say(f"  L-inf Vm {dv.max():.3e} mV; over the {plateau.mean():.3f} activated more than "
        f"10 ms before the end {dv[plateau].max():.3e} mV")
    assert on.mean() > 0.3
    assert dact.max() <= DT * 1.0000001
    assert (last >= end - DT * 1.0000001).all()
    assert dv[plateau].max() < 1.0


def test_mixed_state_roundtrip_and_drug_scales():
    say("mixed: state round trip and per-cell drug scales")
    a = W.brick(12, 10, 8)
    stim = [W.box_stimulus(xMax=3, nx=12, ny=10, nz=8)]
    rng = np.random.default_rng(2)
    # drug scales come per slab / request: a handful of (s_Kr, s_CaL, s_Na) sets
    sets = rng.uniform(0.2, 1.0, (6, 3))
    scales = sets[rng.integers(0, 6, a.ncell)]
    a.celltype[:] = rng.integers(0, 3, a.ncell)
    ref = run(a, stim, "fp64_fused", 6000, scales)
    mix = run(a, stim, "mixed_fused", 6000, scales)
    dact = np.abs(mix["act"] - ref["act"])
    say(f"  activation max diff {dact.max():.4f} ms; Vm L-inf {np.abs(mix['vm'] - ref['vm']).max():.3e}")
    assert dact.max() <= DT * 1.0000001
    # state export/import keeps the mixed state: 1 more step from an
    # imported copy equals 1 more step of the original
    s1 = hip.Simulation(a, W.CELLTYPES_EMI, DT, W.DIFFUSION_SCALE, stim, mode="mixed_fused",
                        scales=scales)
    s1.run(300)
    st, vm = s1.state(), s1.vm()
    s2 = hip.Simulation(a, W.CELLTYPES_EMI, DT, W.DIFFUSION_SCALE, stim, mode="mixed_fused",
                        scales=scales)
    s2.set_state(st, vm)
    np.testing.assert_allclose(s2.state(), st, rtol=1e-6, atol=1e-9)
    assert np.array_equal(s2.vm(), vm)
    s1.close()
    s2.close()


@pytest.mark.parametrize("dims", [(16, 16, 16), (65, 8, 8), (127, 8, 8)])
def test_mixed_state_runs_at_their_edges(dims):
    """The mixed engine keeps its states in per-slab runs of 4096 cells, the
    last run as wide as the cells left: slabs of exactly one run (4096
    cells), one run and a 64-cell remainder (4160), and one run and a 4032-
    cell remainder (8128) activate within one time step of FP64, and a state
    exported and imported mid-run continues as the original does.  The
    import goes to a simulation without the stimulus: the original's pulse
    ended at 1 ms, and the new one's clock starts at 0."""
    nx, ny, nz = dims
    say(f"mixed state runs: brick {nx}x{ny}x{nz} ({nx * ny * nz} cells)")
    a = W.brick(nx, ny, nz)
    stim = [W.box_stimulus(xMax=3, nx=nx, ny=ny, nz=nz)]
    nsteps = 40 * nx
    ref = run(a, stim, "fp64_fused", nsteps)
    mix = run(a, stim, "mixed_fused", nsteps)
    dact = np.abs(mix["act"] - ref["act"])
    say(f"  activated {np.mean(ref['act'] >= 0):.3f}, activation max diff {dact.max():.4f} ms")
    assert (ref["act"] >= 0).all() and (mix["act"] >= 0).all()
    assert dact.max() <= DT * 1.0000001
    s1 = hip.Simulation(a, W.CELLTYPES_EMI, DT, W.DIFFUSION_SCALE, stim, mode="mixed_fused")
    s1.run(100)
    s2 = hip.Simulation(a, W.CELLTYPES_EMI, DT, W.DIFFUSION_SCALE, [], mode="mixed_fused")
    s2.set_state(s1.state(), s1.vm())
    s1.run(100)
    s2.run(100)
    dv = np.abs(s1.vm() - s2.vm()).max()
    say(f"  continued after import: max |dVm| {dv:.3e} mV")
    # K_i is held as an FP32 offset that the dVK_i field round-trips through FP64
    assert dv <= 1e-6
    np.testing.assert_allclose(s2.state(), s1.state(), rtol=1e-5, atol=1e-9)
    s1.close()
    s2.close()
