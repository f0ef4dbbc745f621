"""The Niederer benchmark (dx = 0.2 mm, dt = 0.01 ms, 60 ms) on the engine.

FP64 with the reference's rounding reproduces the oracle's activation times
exactly (the committed fixture, regenerated and checked by
test_oracle_niederer); the FMA and mixed forms are within one time step of
them; every form is within 0.1 ms (+ one step) of the published table
(test/tests/NiedererTest/tools/table.tex:17).
"""
import json
import os
import sys

import numpy as np
import pytest

from cardioid_amd import hip, workloads as W

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
HERE = os.path.dirname(os.path.abspath(__file__))
DT = 0.01


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    hip.lib()
    yield


@pytest.mark.parametrize("mode", ["fp64_fused", "fp64_fma", "mixed_fused", "fp64_unfused"])
def test_niederer_on_the_engine(mode):
    with open(os.path.join(HERE, "golden", "niederer_dx0.2_dt0.01.json")) as fh:
        fx = json.load(fh)
    a, stim, idx = W.niederer(0.2)
    sim = hip.Simulation(a, ["epiCellML"], DT, W.DIFFUSION_SCALE, stim, mode=mode)
    sim.run(6000)
    act = sim.activation()[0]
    sim.close()
    got = np.array([act[i] for i in idx])
    want = np.array(fx["tact0_points"])
    pub = np.array(W.NIEDERER_TABLE[(0.2, 0.01)])
    print(f"{mode}: {dict(zip(W.NIEDERER_POINTS, np.round(got, 3)))}, vs oracle max "
          f"{np.abs(got - want).max():.4f} ms, vs published max {np.abs(got - pub).max():.3f} ms",
          flush=True)
    sys.stdout.flush()
    assert (act >= 0).all()
    if mode in ("fp64_fused", "fp64_unfused"):
        assert np.array_equal(got, want)
        assert float(act.sum()) == fx["tact0_sum"]
    else:
        assert np.abs(got - want).max() <= DT * 1.0000001
    assert np.abs(got - pub).max() <= 0.1 + DT * 1.0000001
