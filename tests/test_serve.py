"""Serving loop logic on the CPU, with a stand-in engine that counts steps.

Checks: arrival-order admission under an arena limit, every request run for
exactly its own number of steps, results collected before its room is
reused, latency = completion - arrival, and an oversized request rejected.
"""
import numpy as np
import pytest

from cardioid_amd import hip, serve, workloads as W


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(dt, 1e-6)


class FakeBatch:
    """Arena of `cap` cells; each run(k) costs k * (active cells) * 1e-9 s."""

    def __init__(self, cap, clock):
        self.cap, self.clock = cap, clock
        self.slabs = {}
        self.next = 0
        self.log = []

    def add(self, slab):
        used = sum(s["n"] for s in self.slabs.values())
        if used + slab.ncell > self.cap:
            raise hip.ArenaFull("full")
        sid = self.next
        self.next += 1
        self.slabs[sid] = {"n": slab.ncell, "steps": 0}
        self.log.append(("add", sid, slab.ncell))
        return sid

    def remove(self, sid):
        self.log.append(("remove", sid, self.slabs[sid]["steps"]))
        del self.slabs[sid]

    def run(self, k):
        for s in self.slabs.values():
            s["steps"] += k
        self.clock.t += k * sum(s["n"] for s in self.slabs.values()) * 1e-9

    def sync(self):
        pass

    def activation(self, sid):
        n = self.slabs[sid]["n"]
        v = np.full(n, float(self.slabs[sid]["steps"]))
        return v, v, v

    def t90(self, sid):
        return np.full(self.slabs[sid]["n"], 2.0 * self.slabs[sid]["steps"])


class FakeSlab:
    def __init__(self, n):
        self.ncell = n


def fake_make_slab(req):
    n = int(np.prod(req.dims))

    class A:
        ncell = n
    return A(), FakeSlab(n)


def test_stream_admission_exact_steps_and_latency():
    reqs = W.request_stream(30, rate_hz=20.0, seed=11)
    clock = FakeClock()
    cap = 3 * 128 ** 3
    b = FakeBatch(cap, clock)
    st = serve.Server(b, dt=0.02, chunk=50, make_slab=fake_make_slab).serve(
        reqs, clock=clock, sleep=clock.sleep)
    assert len(st.tickets) == 30
    by_rid = {t.rid: t for t in st.tickets}
    order = sorted(range(30), key=lambda k: reqs[k].arrival_s)
    for k, r in enumerate(sorted(reqs, key=lambda r: r.arrival_s)):
        t = by_rid[k]
        want = int(round(r.duration_ms / 0.02))
        assert t.nsteps == want
        # the engine ran the slab for exactly its steps: activation (fake) = steps
        assert (t.tact == want).all() and (t.apd90 == want).all()
        assert t.done >= t.admitted >= t.arrival and t.latency > 0
    # admissions in arrival order, never above the arena
    adds = [e for e in b.log if e[0] == "add"]
    assert [e[2] for e in adds] == [int(np.prod(by_rid[k].request.dims)) for k in range(30)]
    live = 0
    for e in b.log:
        live += e[2] if e[0] == "add" else 0
        if e[0] == "remove":
            live -= [a[2] for a in adds if a[1] == e[1]][0]
        assert live <= cap
    s = st.summary()
    assert s["p99_latency_s"] >= s["p50_latency_s"] > 0 and s["max_active"] >= 2
    assert s["cell_steps"] == sum(t.ncell * t.nsteps for t in st.tickets)


def test_request_that_never_fits_is_rejected():
    reqs = [W.Request(0.0, (64, 64, 64), 10.0, np.ones(3), 0.0)]
    clock = FakeClock()
    b = FakeBatch(1000, clock)
    with pytest.raises(RuntimeError, match="never fit"):
        serve.Server(b, make_slab=fake_make_slab).serve(reqs, clock=clock, sleep=clock.sleep)


def test_arena_for_budget():
    cells, box = serve.arena_for_budget(16 * 2 ** 30)
    assert cells % 64 == 0
    per = cells * serve.MIXED_BYTES_PER_CELL + box * serve.BYTES_PER_BOX_POINT
    assert per <= 16 * 2 ** 30 and per > 0.99 * 16 * 2 ** 30 - 64 * 300


def test_split_x_partitions_cells_and_names_neighbour_planes():
    """The x-split of an anatomy (one GPU's share of a sharded run) owns every
    cell once, and each share's remote cells are exactly the other shares'
    tissue on the planes next to its own, inside its padded box."""
    a, _ = W.ellipsoid_shell(30, 26, 40, seed=5)
    shares = W.split_x(a, 4)
    assert len(shares) == 4
    owned = np.concatenate([s.anatomy.gid for s in shares])
    assert np.array_equal(np.sort(owned), np.sort(a.gid))
    sizes = [s.anatomy.ncell for s in shares]
    assert max(sizes) < 1.5 * a.ncell / 4
    x, y, z = a.tuples()
    for k, s in enumerate(shares):
        lo, hi = s.x_range
        sx, sy, sz = s.anatomy.tuples()
        rx = s.remote_gid % a.nx
        assert set(np.unique(rx)) <= {lo - 1, hi}
        # every cell on the neighbour planes within the padded y/z range is remote
        box = ((y >= sy.min() - 1) & (y <= sy.max() + 1) & (z >= sz.min() - 1)
               & (z <= sz.max() + 1) & ((x == lo - 1) | (x == hi))
               & (x >= sx.min() - 1) & (x <= sx.max() + 1))
        assert np.array_equal(np.sort(a.gid[box]), np.sort(s.remote_gid))
        # remote cells belong to the adjacent shares
        nb = set()
        for j in (k - 1, k + 1):
            if 0 <= j < len(shares):
                nb |= set(shares[j].anatomy.gid.tolist())
        assert set(s.remote_gid.tolist()) <= nb


def test_active_cell_cap_keeps_order_and_lowers_the_median():
    """With a cap on cells in flight, an overloaded stream is served in
    arrival order with the cap respected (one request alone may exceed it),
    every request still runs its exact steps, and, the engine's cost being
    proportional to the cells in flight, the median latency drops while the
    total time stays the same."""
    reqs = W.request_stream(30, rate_hz=20.0, seed=11)
    res = {}
    for cap in (None, 2 * 128 ** 3):
        clock = FakeClock()
        b = FakeBatch(64 * 128 ** 3, clock)
        st = serve.Server(b, dt=0.02, chunk=50, make_slab=fake_make_slab,
                          max_active_cells=cap).serve(reqs, clock=clock, sleep=clock.sleep)
        assert len(st.tickets) == 30
        for t in st.tickets:
            assert (t.tact == t.nsteps).all()
        live, peak = {}, 0
        for e in b.log:
            if e[0] == "add":
                live[e[1]] = e[2]
                if cap is not None and len(live) > 1:
                    assert sum(live.values()) <= cap
            else:
                del live[e[1]]
        res[cap] = st.summary()
    free, capped = res[None], res[2 * 128 ** 3]
    assert capped["p50_latency_s"] < 0.8 * free["p50_latency_s"]
    assert capped["wall_s"] <= 1.05 * free["wall_s"]


def test_least_work_first_admission():
    """order="shortest" admits the waiting request with the least cells x
    steps first: every request still completes with its exact steps, the
    cap still holds, and the median latency of an overloaded stream drops
    below arrival order's."""
    reqs = W.request_stream(30, rate_hz=20.0, seed=11)
    res = {}
    for order in ("arrival", "shortest"):
        clock = FakeClock()
        b = FakeBatch(64 * 128 ** 3, clock)
        st = serve.Server(b, dt=0.02, chunk=50, make_slab=fake_make_slab,
                          max_active_cells=2 * 128 ** 3, order=order).serve(
            reqs, clock=clock, sleep=clock.sleep)
        assert len(st.tickets) == 30
        assert all((t.tact == t.nsteps).all() for t in st.tickets)
        res[order] = st.summary()
    assert res["shortest"]["p50_latency_s"] < res["arrival"]["p50_latency_s"]
    with pytest.raises(ValueError):
        serve.Server(None, order="random")
