"""Checkpoint / restart (elec/checkpointIO.cc writeCheckpoint, readCheckpoint)
of the engine on an MI355X: a run that writes a checkpoint at step 150, a
fresh engine restarted from it and run 150 more steps, and an uninterrupted
300-step run end in the same state.

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


@pytest.mark.parametrize("mode,tol", [("fp64_fused", 0.0), ("mixed_fused", 1e-6)])
def test_batch_slab_evicted_and_resumed(tmp_path, mode, tol):
    """A slab of the batched engine written out at step 150, removed while
    the other slab runs on, added back into a different arena slot and read
    back, then run to step 300, ends where an uninterrupted run ends (bit for
    bit in FP64; the mixed form as in the single-engine test)."""
    cases = W.batched_slabs(n=2, seed=7, lo=(24, 24, 24), hi=(40, 40, 40))
    slabs = [hip.Slab(c.anatomy, ["endoCellML"], W.DIFFUSION_SCALE, c.stimuli, scales=c.scales)
             for c in cases]
    ext = [s.extent() for s in slabs]
    cells, boxes = sum(e[0] for e in ext), sum(e[1] for e in ext)
    say(f"batch slab eviction and resume: {mode}, slabs of {[s.ncell for s in slabs]} cells")
    b = hip.Batch(DT, cells, boxes, slab_capacity=2, mode=mode)
    ids = [b.add(s) for s in slabs]
    b.run(300)
    want = b.vm(ids[1])
    b.close()

    b = hip.Batch(DT, cells, boxes, slab_capacity=2, mode=mode)
    ids = [b.add(s) for s in slabs]
    b.set_graph(10)
    b.run(150)
    path = b.write_checkpoint(ids[1], str(tmp_path), ascii=False)
    b.remove(ids[1])
    b.run(50)
    again = b.add(slabs[1])
    b.read_checkpoint(again, path)
    assert b.slab_info(again)["loop"] == 150
    b.run(150)
    assert b.slab_info(again)["loop"] == 300
    got = b.vm(again)
    b.close()
    dv = np.abs(got - want).max()
    say(f"  max |dVm| {dv:.3e} mV")
    assert want.max() > 0.0
    if tol == 0.0:
        assert np.array_equal(got, want)
    else:
        assert dv <= tol
