"""Checkpoint / restart (elec/checkpointIO.cc writeCheckpoint, readCheckpoint)
of the engine on an MI355X: a run that writes a checkpoint at step 150, a
fresh engine restarted from it and run 150 more steps, and an uninterrupted
300-step run end in the same state; a restarted run equals the oracle
restarted from the same file (states, clock, fresh sensors), activation
times and APD90 included; an evicted and resumed batch slab keeps its
sensors.

Bar: bitwise for the FP64 modes from a FIXRECORDBINARY file (and across
engines: a checkpoint of the fused engine restarts the reference loop);
FIXRECORDASCII keeps 14 significant digits, so Vm within 1e-9 mV; the mixed
engine stores K_i as an FP32 offset that the dVK_i field round-trips through
FP64 arithmetic, so Vm within 1e-6 mV.
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
    hip.lib()
    yield


def _case():
    a = W.brick(12, 10, 8)
    stim = [W.box_stimulus(xMax=3, nx=12, ny=10, nz=8, duration=1.0)]
    return a, stim


def _sim(mode):
    a, stim = _case()
    return hip.Simulation(a, W.CELLTYPES_EMI[:1], DT, W.DIFFUSION_SCALE, stim, mode=mode)


@pytest.mark.parametrize("mode,restart_mode,ascii,tol", [
    ("fp64_fused", "fp64_fused", False, 0.0),
    ("fp64_fma", "fp64_fma", False, 0.0),
    ("fp64_unfused", "fp64_unfused", False, 0.0),
    ("fp64_fused", "fp64_unfused", False, 0.0),
    ("fp64_fused", "fp64_fused", True, 1e-9),
    ("mixed_fused", "mixed_fused", False, 1e-6),
])
def test_restart_continues_the_run(tmp_path, mode, restart_mode, ascii, tol):
    say(f"checkpoint {mode} -> restart {restart_mode}, ascii={ascii}")
    whole = _sim(mode)
    whole.run(300)
    vm_want, st_want = whole.vm(), whole.state()
    whole.close()

    first = _sim(mode)
    first.run(150)
    path = first.write_checkpoint(str(tmp_path), ascii=ascii)
    vm_mid = first.vm()
    first.close()
    assert path.endswith("snapshot.000000000150/state")

    second = _sim(restart_mode)
    second.read_checkpoint(path)
    assert second.loop == 150
    assert abs(second.time - 150 * DT) < 1e-6
    if not ascii and mode == restart_mode and mode != "mixed_fused":
        np.testing.assert_array_equal(second.vm(), vm_mid)
    second.run(150)
    vm, st = second.vm(), second.state()
    second.close()
    dv = np.abs(vm - vm_want).max()
    say(f"  max |dVm| {dv:.3e} mV")
    assert np.isfinite(vm).all() and vm_want.max() > 0.0     # the wave is running
    if tol == 0.0:
        np.testing.assert_array_equal(vm, vm_want)
        np.testing.assert_array_equal(st, st_want)
    else:
        assert dv <= tol


@pytest.mark.parametrize("mode", ["fp64_fused", "fp64_unfused", "fp64_fma"])
def test_restart_matches_the_oracle_restarted_from_the_same_file(tmp_path, mode):
    """The oracle in the loop.  HIP runs 150 steps and writes a binary
    checkpoint; a fresh HIP simulation restarts from it, and the oracle,
    after its own 150 steps, restarts from the same HIP-written file (read
    back through cdh_checkpoint_read) with fresh sensors, as the reference
    restarts (its sensors are created anew).  Both run on to step 17000
    (340 ms, past repolarisation): Vm within 1e-9 relative, tact0 / tact /
    trec and t90 identical.  At the restart (3 ms) part of the brick is
    above the threshold: the fresh sensors record those cells at the first
    step after the restart, in both."""
    import oracle_lib as OL
    a, stim = _case()
    say(f"restart vs oracle: {mode}, brick 12x10x8, checkpoint at step 150, run to 17000")
    first = hip.Simulation(a, W.CELLTYPES_EMI[:1], DT, W.DIFFUSION_SCALE, stim, mode=mode,
                           sensors=3)
    first.run(150)
    path = first.write_checkpoint(str(tmp_path), ascii=False)
    first.close()
    vm0, st0, loop, time = hip.read_checkpoint(path, a.gid)
    assert loop == 150 and abs(time - 3.0) < 1e-6
    above = vm0 > -40.0
    say(f"  cells above -40 mV at the restart: {above.mean():.3f}")
    assert 0.0 < above.mean() < 1.0

    sim = hip.Simulation(a, W.CELLTYPES_EMI[:1], DT, W.DIFFUSION_SCALE, stim, mode=mode,
                         sensors=3)
    sim.read_checkpoint(path)
    ref = OL.OracleSim(a, W.CELLTYPES_EMI[:1], DT, W.DIFFUSION_SCALE, stim)
    ref.run(150)
    ref.restart(st0, vm0, loop, time)
    for n in (1, 16849):
        sim.run(n)
        ref.run(n)
        if n == 1:
            t1 = sim.time
            act = sim.activation()[1]
            # the fresh sensors take the cells above the level at once
            assert (act[above] == t1).all()
    assert sim.time == ref.time and sim.loop == 17000
    vm, want = sim.vm(), ref.vm()
    e = np.abs(vm - want) / np.maximum(np.abs(want), 1.0)
    say(f"  Vm max rel {e.max():.3e}")
    assert e.max() < 1e-9
    got, exp = sim.activation(), ref.activation()
    for k, nm in enumerate(("tact0", "tact", "trec")):
        same = got[k] == exp[k]
        say(f"  {nm} identical {same.mean():.4f}, set {np.mean(exp[k] >= 0):.3f}")
        assert same.all(), nm
    t90, t90w = sim.t90(), ref.t90()
    say(f"  t90 identical {np.mean(t90 == t90w):.4f}, reached {np.mean(t90w > 0):.3f}")
    assert np.array_equal(t90, t90w) and (t90w > 0).all()
    sim.close()
    ref.close()


@pytest.mark.parametrize("mode,tol", [("fp64_fused", 0.0), ("mixed_fused", 1e-2)])
def test_batch_slab_evicted_and_resumed(tmp_path, mode, tol):
    """A slab of the batched engine written out at step 150 with its sensors
    and exact clock, removed while the other slab runs on, added back into a
    different arena slot, read back and its sensors restored, then run to
    step 16000 (320 ms), ends where an uninterrupted run ends: Vm,
    activation times and t90 bit for bit in FP64.  The mixed engine's K_i
    offset round-trips through the FP64 dVK_i field (a last-bit change), so
    there activation lands within one step, t90 within two and Vm within
    0.01 mV."""
    cases = W.batched_slabs(n=2, seed=7, lo=(24, 24, 24), hi=(40, 40, 40))
    slabs = [hip.Slab(c.anatomy, ["endoCellML"], W.DIFFUSION_SCALE, c.stimuli, scales=c.scales)
             for c in cases]
    ext = [s.extent() for s in slabs]
    cells, boxes = sum(e[0] for e in ext), sum(e[1] for e in ext)
    say(f"batch slab eviction and resume: {mode}, slabs of {[s.ncell for s in slabs]} cells")
    total, at = 16000, 150
    b = hip.Batch(DT, cells, boxes, slab_capacity=2, mode=mode, sensors=3)
    ids = [b.add(s) for s in slabs]
    b.run(at)
    mid_act = b.activation(ids[1])[1]
    b.run(total - at)
    want = b.vm(ids[1])
    want_act, want_t90 = b.activation(ids[1]), b.t90(ids[1])
    want_time = b.slab_info(ids[1])["time"]
    b.close()

    b = hip.Batch(DT, cells, boxes, slab_capacity=2, mode=mode, sensors=3)
    ids = [b.add(s) for s in slabs]
    b.set_graph(10)
    b.run(at)
    path = b.write_checkpoint(ids[1], str(tmp_path), ascii=False)
    rows, t_evict, l_evict = b.sensors(ids[1])
    assert l_evict == at and (rows[1] >= 0).any()
    b.remove(ids[1])
    b.run(50)
    again = b.add(slabs[1])
    b.read_checkpoint(again, path)
    assert b.slab_info(again)["loop"] == at
    # the checkpoint restarts fresh sensors (the reference's restart) ...
    assert (b.activation(again)[1] < 0).all()
    # ... and the saved sensors and exact clock resume the served request
    b.set_sensors(again, rows, t_evict, l_evict)
    np.testing.assert_array_equal(b.activation(again)[1], mid_act)
    b.run(total - at)
    assert b.slab_info(again)["loop"] == total
    assert b.slab_info(again)["time"] == want_time
    got = b.vm(again)
    got_act, got_t90 = b.activation(again), b.t90(again)
    b.close()
    dv = np.abs(got - want).max()
    say(f"  max |dVm| {dv:.3e} mV; t90 reached {np.mean(want_t90 > 0):.3f}")
    assert (want_t90 > 0).mean() > 0.5
    if tol == 0.0:
        assert np.array_equal(got, want)
        for k in range(3):
            assert np.array_equal(got_act[k], want_act[k])
        assert np.array_equal(got_t90, want_t90)
    else:
        assert dv <= tol
        assert np.abs(got_act[1] - want_act[1]).max() <= DT * 1.0000001
        ok = (want_t90 > 0) & (got_t90 > 0)
        assert np.array_equal(want_t90 > 0, got_t90 > 0)
        assert np.abs(got_t90 - want_t90)[ok].max() <= 2 * DT
