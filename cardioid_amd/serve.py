"""Serving loop: simulation requests admitted into one batched engine.

Each request is a tissue slab to simulate for a given time; its result is
the activation-time map and the APD90 of every cell.  Requests wait in
arrival order until the engine's arena (sized once from an HBM budget) has
room; admitted requests keep their state on the device and advance together
in the engine's launches (continuous batching); a request leaves the batch
at exactly its own step count and frees its room for the next.  While the
device runs a chunk of steps, the host prepares the next arrivals.

There is no serving loop in the reference; one request is what one Cardioid
run (elec/simulationLoop.cc:150-243 on its anatomy, with the
ActivationAndRecovery sensor) computes.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field
from typing import Callable, Deque, Dict, List, Optional

import numpy as np

from . import hip, workloads as W

# device bytes the mixed engine holds per arena unit (engine_host.hip)
MIXED_BYTES_PER_CELL = (4 + 4            # meta, box index
                        + 17 * 4 + 2 * 8  # FP32 and FP64 states
                        + 18 * 4          # stencil weights
                        + 3 * 8           # activation sensors
                        + 1 + 2 * 8)      # APD90 sensor
BYTES_PER_BOX_POINT = 2 * 8               # double-buffered Vm box


def arena_for_budget(hbm_bytes: int, box_per_cell: float = 1.6):
    """(cells, box points) of an arena that fits `hbm_bytes`, with
    `box_per_cell` box points per cell (a slab's padded box over its tissue
    cells; 1.4-1.6 for the request shapes here)."""
    per_cell = MIXED_BYTES_PER_CELL + box_per_cell * BYTES_PER_BOX_POINT
    cells = int(hbm_bytes // per_cell) // 64 * 64
    return cells, int(cells * box_per_cell)


@dataclass
class Ticket:
    rid: int
    request: W.Request
    nsteps: int
    arrival: float                      # wall clock (perf_counter) of arrival
    admitted: float = 0.0
    done: float = 0.0
    slab: int = -1
    remaining: int = 0
    ncell: int = 0
    prepared: Optional[tuple] = None    # (anatomy, slab description), built once
    tact: Optional[np.ndarray] = None
    apd90: Optional[np.ndarray] = None

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def request_slab(req: W.Request, diffusion_scale: float = W.DIFFUSION_SCALE):
    """The tissue slab of a request: an all-tissue endocardial slab with the
    request's fibre angle and drug scales, stimulated at x < 4."""
    nx, ny, nz = req.dims
    a = W.fibre_slab(nx, ny, nz, req.phi_deg)
    stim = [W.box_stimulus(xMax=4, nx=nx, ny=ny, nz=nz)]
    return a, hip.Slab(a, ["endoCellML"], diffusion_scale, stim, scales=req.scales)


@dataclass
class ServeStats:
    tickets: List[Ticket] = field(default_factory=list)
    wall_s: float = 0.0
    chunks: int = 0
    max_active: int = 0

    def summary(self) -> Dict[str, float]:
        lat = np.array([t.latency for t in self.tickets])
        wait = np.array([t.admitted - t.arrival for t in self.tickets])
        cs = sum(t.ncell * t.nsteps for t in self.tickets)
        return {"requests": len(self.tickets), "p50_latency_s": float(np.percentile(lat, 50)),
                "p99_latency_s": float(np.percentile(lat, 99)),
                "mean_queue_wait_s": float(wait.mean()), "wall_s": self.wall_s,
                "cell_steps": float(cs), "gcell_steps_per_s": cs / self.wall_s / 1e9,
                "max_active": self.max_active, "chunks": self.chunks}


class Server:
    """Continuous-batching loop over one engine (hip.Batch or an object with
    the same methods)."""

    def __init__(self, batch, dt: float = 0.02, chunk: int = 100,
                 make_slab: Callable = request_slab, max_active_cells: Optional[int] = None,
                 order: str = "arrival"):
        """max_active_cells: admit a request only while the cells in flight
        stay within this many (one request is always admitted when none is in
        flight).  Enough cells to fill the GPU keep its throughput, and
        requests then finish in arrival order instead of all sharing the GPU
        to the end, which lowers the median latency of an overloaded stream.
        None: the arena alone limits admission.

        order: "arrival" admits the waiting requests first come, first served;
        "shortest" admits the one with the least work (cells x steps) first,
        which lowers the median latency further at the expense of the longest
        requests."""
        if order not in ("arrival", "shortest"):
            raise ValueError("order must be 'arrival' or 'shortest'")
        self.order = order
        self.batch = batch
        self.dt = dt
        self.chunk = chunk
        self.make_slab = make_slab
        self.max_active_cells = max_active_cells

    def serve(self, requests: List[W.Request], clock=time.perf_counter,
              sleep=time.sleep) -> ServeStats:
        """Runs the stream: request k arrives `arrival_s` after the start."""
        pending: Deque[W.Request] = deque(sorted(requests, key=lambda r: r.arrival_s))
        queue: Deque[Ticket] = deque()
        active: Dict[int, Ticket] = {}
        stats = ServeStats()
        t0 = clock()
        nid = [0]

        def arrivals():
            now = clock()
            while pending and t0 + pending[0].arrival_s <= now:
                r = pending.popleft()
                queue.append(Ticket(nid[0], r, int(round(r.duration_ms / self.dt)),
                                    t0 + r.arrival_s))
                nid[0] += 1

        def work(t):
            return int(np.prod(t.request.dims)) * t.nsteps

        def admit():
            # in arrival (or least-work) order while the arena has room; slabs
            # added while a chunk runs join at the next chunk (the engine
            # orders the uploads)
            while queue:
                if self.order == "shortest" and len(queue) > 1:
                    k = min(range(len(queue)), key=lambda j: (work(queue[j]), queue[j].rid))
                    if k:
                        t = queue[k]
                        del queue[k]
                        queue.appendleft(t)
                t = queue[0]
                if t.prepared is None:
                    t.prepared = self.make_slab(t.request)
                a, slab = t.prepared
                if (self.max_active_cells and active and
                        sum(x.ncell for x in active.values()) + a.ncell > self.max_active_cells):
                    return
                try:
                    t.slab = self.batch.add(slab)
                except hip.ArenaFull:
                    if not active:
                        raise RuntimeError(f"request {t.rid} ({t.request.dims}) can never fit "
                                           "the arena")
                    return
                queue.popleft()
                t.prepared = None
                t.admitted = clock()
                t.remaining = t.nsteps
                t.ncell = a.ncell
                active[t.slab] = t

        while pending or queue or active:
            arrivals()
            admit()
            if not active:
                if pending:
                    sleep(max(0.0, t0 + pending[0].arrival_s - clock()))
                continue
            stats.max_active = max(stats.max_active, len(active))
            running = dict(active)
            k = min(self.chunk, min(t.remaining for t in running.values()))
            self.batch.run(k)
            stats.chunks += 1
            # host work while the device runs
            arrivals()
            admit()
            self.batch.sync()
            for sid, t in running.items():
                t.remaining -= k
                if t.remaining == 0:
                    tact = self.batch.activation(sid)[1]
                    t90 = self.batch.t90(sid)
                    t.tact = tact
                    t.apd90 = np.where((t90 >= 0) & (tact >= 0), t90 - tact, -1.0)
                    self.batch.remove(sid)
                    t.done = clock()
                    stats.tickets.append(t)
                    del active[sid]
        stats.wall_s = clock() - t0
        stats.tickets.sort(key=lambda t: t.rid)
        return stats


def run_alone(req: W.Request, dt: float = 0.02, make_slab: Callable = request_slab):
    """One request in an engine of its own: (activation map, APD90)."""
    a, slab = make_slab(req)
    cells, box, _ = slab.extent()
    b = hip.Batch(dt, cells, box, slab_capacity=1, mode="mixed_fused", sensors=3)
    sid = b.add(slab)
    b.run(int(round(req.duration_ms / dt)))
    tact = b.activation(sid)[1]
    t90 = b.t90(sid)
    b.close()
    return tact, np.where((t90 >= 0) & (tact >= 0), t90 - tact, -1.0)


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--requests", type=int, default=40)
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--hbm-gb", type=float, default=16.0)
    ap.add_argument("--chunk", type=int, default=100)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--graph", type=int, default=10)
    ap.add_argument("--max-active-cells", type=int, default=0,
                    help="cells in flight at most (0: the HBM budget alone)")
    ap.add_argument("--order", choices=("arrival", "shortest"), default="arrival")
    args = ap.parse_args(argv)
    reqs = W.request_stream(args.requests, rate_hz=args.rate, seed=args.seed)
    cells, box = arena_for_budget(int(args.hbm_gb * 2 ** 30))
    b = hip.Batch(0.02, cells, box, slab_capacity=1024, mode="mixed_fused", sensors=3)
    b.set_graph(args.graph)
    stats = Server(b, chunk=args.chunk, max_active_cells=args.max_active_cells or None,
                   order=args.order).serve(reqs)
    out = stats.summary()
    out.update(hbm_budget_gb=args.hbm_gb, arena_cells=cells, arena_box=box,
               rate_hz=args.rate, max_active_cells=args.max_active_cells, order=args.order,
               info=b.info())
    b.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
