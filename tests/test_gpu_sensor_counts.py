"""Device counts of recorded activation and recovery times, on an MI355X.

The step kernels (and the drop-in sensor object's kernel) count the cells
whose tact0 (first Vm > 0), tact (threshold up-crossing) and trec (the
down-crossing after it) they record, with one wavefront ballot and atomic
per kind.  Each count must equal the number of recorded times in the
arrays, in every mode, through arrivals and departures in a batch, an
autotune (which restores what it runs over), an evicted and resumed slab
(counts set from the restored times) and a restart (fresh sensors).
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


def recorded(act):
    return tuple(int(np.count_nonzero(a >= 0)) for a in act)


@pytest.mark.parametrize("mode", ["fp64_unfused", "fp64_fused", "fp64_fma", "mixed_fused"])
def test_counts_equal_recorded_times_in_every_mode(mode):
    a = W.brick(18, 14, 10)
    stim = [W.box_stimulus(xMax=3, nx=18, ny=14, nz=10)]
    sim = hip.Simulation(a, ["endoCellML"], DT, W.DIFFUSION_SCALE, stim, mode=mode)
    seen = []
    # upstroke, plateau and repolarisation (recovery near 280 ms)
    for n in (100, 100, 100, 14000, 2000):
        sim.run(n)
        got, want = sim.sensor_counts(), recorded(sim.activation())
        say(f"  {mode} loop {sim.loop}: counts {got}, recorded {want}")
        assert got == want
        seen.append(got)
    assert any(0 < c[1] < a.ncell for c in seen[:3])    # caught while it spreads
    assert seen[-1][1] == a.ncell and seen[-1][2] > 0
    sim.close()


def test_batch_counts_through_arrivals_autotune_eviction_and_restart(tmp_path):
    cases = W.batched_slabs(n=3, seed=5, lo=(16, 16, 16), hi=(28, 28, 28))
    slabs = [hip.Slab(c.anatomy, ["endoCellML"], W.DIFFUSION_SCALE, c.stimuli, scales=c.scales)
             for c in cases]
    ext = [s.extent() for s in slabs]
    b = hip.Batch(DT, sum(e[0] for e in ext), sum(e[1] for e in ext), slab_capacity=3,
                  mode="fp64_fused", sensors=3)

    def check(ids, what):
        for sid in ids:
            got, want = b.sensor_counts(sid), recorded(b.activation(sid))
            say(f"  {what}, slab {sid}: counts {got}, recorded {want}")
            assert got == want

    ids = [b.add(slabs[0]), b.add(slabs[1])]
    b.run(300)
    check(ids, "two slabs, 300 steps")
    ids.append(b.add(slabs[2]))               # arrives while the others run
    b.run(200)
    check(ids, "a third arrives")
    b.autotune(4)
    b.run(100)
    check(ids, "after an autotune")
    # evict slab 1 with its sensors, bring it back into another arena slot
    (tmp_path / "ev").mkdir()
    (tmp_path / "rs").mkdir()
    state = b.write_checkpoint(ids[1], str(tmp_path / "ev"), ascii=False)
    rows, t, loop = b.sensors(ids[1])
    b.remove(ids[1])
    b.run(50)
    back = b.add(slabs[1])
    b.read_checkpoint(back, state)
    b.set_sensors(back, rows, t, loop)
    ids[1] = back
    check(ids, "evicted and resumed")
    b.run(400)
    check(ids, "resumed slab runs on")
    # restart: fresh sensors, counts from zero, then counting again
    state = b.write_checkpoint(ids[0], str(tmp_path / "rs"), ascii=True)
    b.read_checkpoint(ids[0], state)
    assert b.sensor_counts(ids[0]) == (0, 0, 0)
    b.run(100)
    check(ids, "after a restart")
    b.close()
