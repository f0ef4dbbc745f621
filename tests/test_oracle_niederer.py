"""Whole-path anchor: the Niederer et al. (2011) N-version benchmark.

The oracle runs the reference's own benchmark set-up (3 x 7 x 20 mm epi slab,
fibres along z, corner box stimulus, dx = 0.2 mm, dt = 0.01 ms, 60 ms) and
its activation times at the benchmark's nine points are compared with the
table the reference publishes for that resolution
(test/tests/NiedererTest/tools/table.tex:17).  The published run used the
SIMD non-gate variant (fastNonGate = 1) with approximate logs; this path is
the generic non-gate update, so agreement is to a few hundredths of a ms,
not bitwise: the bar is 0.1 ms (0.3 % of the latest point).  The same run
must also reproduce the committed fixture that the GPU test compares with.
"""
import json
import os

import numpy as np
import pytest

from cardioid_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.slow
def test_niederer_benchmark_activation_times():
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    from make_niederer import run
    act, idx = run(nthreads=min(8, os.cpu_count() or 1))
    got = np.array([act[i] for i in idx])
    pub = np.array(W.NIEDERER_TABLE[(0.2, 0.01)])
    d = got - pub
    print("Niederer dx 0.2 dt 0.01:", dict(zip(W.NIEDERER_POINTS, np.round(got, 3))),
          "max |diff| vs published", np.abs(d).max())
    assert np.abs(d).max() <= 0.1
    assert (act >= 0).all()
    with open(os.path.join(HERE, "golden", "niederer_dx0.2_dt0.01.json")) as fh:
        fx = json.load(fh)
    assert np.array_equal(got, np.array(fx["tact0_points"]))
    assert int((act >= 0).sum()) == fx["activated"]
    assert float(act[act >= 0].sum()) == fx["tact0_sum"]
