"""Serving loop on an MI355X: requests that arrive while others run are
admitted as room frees up, advance in the same launches, and each result
(activation map and APD90) is bit for bit the one the request gets alone."""
import sys

import numpy as np
import pytest

from cardioid_amd import hip, serve, workloads as W

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    hip.lib()
    yield


def test_results_alone_equal_results_in_the_mix():
    reqs = W.request_stream(8, rate_hz=200.0, seed=5, lo=8, hi=20, dur_ms=(40.0, 120.0))
    # room for about three of them at once
    ext = [serve.request_slab(r)[1].extent() for r in reqs]
    cells = 3 * max(e[0] for e in ext)
    box = 3 * max(e[1] for e in ext)
    b = hip.Batch(0.02, cells, box, slab_capacity=16, mode="mixed_fused", sensors=3)
    b.set_graph(10)
    st = serve.Server(b, chunk=64).serve(reqs)
    b.close()
    s = st.summary()
    say(f"  {s['requests']} requests, max {s['max_active']} in flight, {s['chunks']} chunks, "
        f"p50 {s['p50_latency_s']:.3f} s, p99 {s['p99_latency_s']:.3f} s")
    assert s["requests"] == 8 and s["max_active"] >= 2
    for t in st.tickets:
        tact, apd = serve.run_alone(t.request)
        same = np.array_equal(tact, t.tact) and np.array_equal(apd, t.apd90)
        say(f"  request {t.rid}: {t.request.dims}, {t.nsteps} steps, activated "
            f"{(tact >= 0).mean():.2f}, identical alone {same}")
        assert same


def test_cfg4_stream_alone_equals_in_the_mix():
    """cfg4's requests at their own configuration: the first 10 requests of
    the Poisson stream at 20 req/s (seed 4), dims log-uniform 32..128 per
    axis, 200-1000 ms each, drug scales U[0, 1], served on a 24 GB budget
    with continuous batching; each request's activation map and APD90 equal,
    bit for bit, those of the same request run alone."""
    reqs = W.request_stream(10, rate_hz=20.0, seed=4)
    cells, box = serve.arena_for_budget(24 * 2 ** 30)
    b = hip.Batch(0.02, cells, box, slab_capacity=64, mode="mixed_fused", sensors=3)
    b.set_graph(10)
    say(f"cfg4: {len(reqs)} requests, dims {[r.dims for r in reqs]}")
    st = serve.Server(b, chunk=100).serve(reqs)
    b.close()
    s = st.summary()
    say(f"  max {s['max_active']} in flight, {s['chunks']} chunks, p50 {s['p50_latency_s']:.2f} s, "
        f"p99 {s['p99_latency_s']:.2f} s, {s['gcell_steps_per_s']:.2f} Gcell-steps/s")
    assert s["requests"] == 10 and s["max_active"] >= 2
    for t in st.tickets:
        tact, apd = serve.run_alone(t.request)
        same = np.array_equal(tact, t.tact) and np.array_equal(apd, t.apd90)
        say(f"  request {t.rid}: {t.request.dims}, {t.nsteps} steps, activated "
            f"{(tact >= 0).mean():.2f}, repolarised {(apd > 0).mean():.2f}, identical alone {same}")
        assert same
