"""The batched engine (cdh_batch_*, through the C ABI) on an MI355X.

- FP64 slabs in one batch each match the oracle run of that slab alone (the
  drop-in bar: Vm within 1e-9 relative after the run, activation identical).
- Every slab's result is the same alone and in a batch, bit for bit, in both
  precisions, with slabs added and removed while others run.
- Steps replayed from captured HIP graphs equal directly launched steps.
- The autotuner leaves the simulation where it found it.
- An anatomy split along x into slabs whose halo planes are exchanged after
  every step (the one-GPU form of the sharded run) runs as the whole anatomy:
  bit for bit in FP64; in the mixed form, whose compact weights depend on the
  tiling, with activation times within 1 dt.
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


def relerr(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1.0)


def small_cases():
    """Three unlike slabs: a brick with a box stimulus, a shell with three
    cell types and point stimuli, and a fibre slab under drug block.
    Returns (anatomy, cell types, stimuli, per-cell scales or None, Slab)."""
    a1 = W.brick(10, 8, 6)
    st1 = [W.box_stimulus(xMax=3, nx=10, ny=8, nz=6)]
    a2, depth = W.ellipsoid_shell(16, 14, 20, seed=2)
    st2 = W.endocardial_point_stimuli(a2, depth, 2, seed=2)
    a3 = W.fibre_slab(9, 7, 5, 30.0)
    sc3 = np.array([0.5, 0.7, 0.9])
    st3 = [W.box_stimulus(xMax=4, nx=9, ny=7, nz=5)]
    return [(a1, ["endoCellML"], st1, None,
             hip.Slab(a1, ["endoCellML"], W.DIFFUSION_SCALE, st1)),
            (a2, W.CELLTYPES_EMI, st2, None,
             hip.Slab(a2, W.CELLTYPES_EMI, W.DIFFUSION_SCALE, st2)),
            (a3, ["endoCellML"], st3, np.tile(sc3, a3.ncell),
             hip.Slab(a3, ["endoCellML"], W.DIFFUSION_SCALE, st3, scales=sc3))]


def capacity(slabs):
    cells = sum(s.extent()[0] for s in slabs)
    box = sum(s.extent()[1] for s in slabs)
    return cells, box


def test_fp64_batch_matches_oracle_per_slab():
    import oracle_lib as OL
    cases = small_cases()
    slabs = [c[4] for c in cases]
    cells, box = capacity(slabs)
    nsteps = 600
    say(f"fp64 batch of {len(slabs)} slabs, {nsteps} steps, graphs of 4 steps")
    b = hip.Batch(DT, cells, box, slab_capacity=8, mode="fp64_fused")
    b.set_graph(4)
    ids = [b.add(s) for s in slabs]
    b.run(nsteps)
    for (a, cts, stims, sc, _), sid in zip(cases, ids):
        ref = OL.OracleSim(a, cts, DT, W.DIFFUSION_SCALE, stims, scales=sc)
        ref.run(nsteps)
        e = relerr(b.vm(sid), ref.vm())
        act, want = b.activation(sid), ref.activation()
        info = b.slab_info(sid)
        say(f"  slab {sid}: {a.ncell} cells, Vm max rel {e.max():.2e}, "
            f"activated {(act[1] >= 0).mean():.3f}, time {info['time']:.4f}")
        assert info["time"] == ref.time and info["loop"] == nsteps
        assert e.max() < 1e-9
        for k in range(3):
            assert np.array_equal(act[k], want[k])
        ref.close()
    b.close()


@pytest.mark.parametrize("mode", ["fp64_fused", "mixed_fused"])
def test_alone_equals_in_batch_with_arrivals_and_departures(mode):
    """Slab C joins after 150 steps while A and B run; B leaves at 300 and D
    takes its place.  Each slab's Vm, states and activation times equal those
    of the same slab run alone for the same number of its own steps."""
    cases = small_cases()
    sa, sb, sc = cases[0][4], cases[1][4], cases[2][4]
    a4 = W.fibre_slab(12, 6, 4, -45.0)
    sd = hip.Slab(a4, ["epiCellML"], W.DIFFUSION_SCALE, [W.box_stimulus(xMax=2, nx=12, ny=6, nz=4)],
                  scales=np.array([0.2, 1.0, 0.6]))
    cells, box = capacity([sa, sb, sc, sd])
    say(f"{mode}: arrivals and departures in one batch")
    b = hip.Batch(DT, cells, box, slab_capacity=4, mode=mode, sensors=3)
    b.set_graph(6)
    ia, ib = b.add(sa), b.add(sb)
    b.run(150)
    ic = b.add(sc)
    b.run(150)
    got_b = (b.vm(ib), b.state(ib), b.activation(ib))
    b.remove(ib)
    idd = b.add(sd)
    b.run(401)
    got = {"a": (b.vm(ia), b.state(ia), b.activation(ia), b.t90(ia), 701),
           "c": (b.vm(ic), b.state(ic), b.activation(ic), b.t90(ic), 551),
           "d": (b.vm(idd), b.state(idd), b.activation(idd), b.t90(idd), 401)}
    b.close()

    def alone(slab, n):
        c, x = capacity([slab])
        one = hip.Batch(DT, c, x, slab_capacity=1, mode=mode, sensors=3)
        s = one.add(slab)
        one.run(n)
        out = (one.vm(s), one.state(s), one.activation(s), one.t90(s))
        one.close()
        return out

    ref_b = alone(sb, 300)
    assert np.array_equal(got_b[0], ref_b[0]) and np.array_equal(got_b[1], ref_b[1])
    for k in range(3):
        assert np.array_equal(got_b[2][k], ref_b[2][k])
    for name, slab in (("a", sa), ("c", sc), ("d", sd)):
        g = got[name]
        r = alone(slab, g[4])
        same = [np.array_equal(g[0], r[0]), np.array_equal(g[1], r[1]),
                all(np.array_equal(g[2][k], r[2][k]) for k in range(3)), np.array_equal(g[3], r[3])]
        say(f"  slab {name}: {g[4]} steps, identical (Vm, states, activation, t90) = {same}")
        assert all(same)


def test_graph_replay_equals_direct_launches():
    cases = small_cases()
    slabs = [c[4] for c in cases]
    cells, box = capacity(slabs)
    out = []
    for graph in (0, 2, 10):
        b = hip.Batch(DT, cells, box, slab_capacity=4, mode="mixed_fused")
        b.set_graph(graph)
        ids = [b.add(s) for s in slabs]
        for chunk in (7, 40, 33, 120):     # odd chunks: direct steps between replays
            b.run(chunk)
            b.sync()
            _ = b.vm(ids[0])
        out.append([np.concatenate([b.vm(i), b.state(i).ravel(), b.activation(i)[1]]) for i in ids])
        b.close()
    for g in (1, 2):
        for k in range(len(slabs)):
            assert np.array_equal(out[0][k], out[g][k])
    say("  graphs of 2 and 10 steps replay the direct launches bit for bit")


@pytest.mark.parametrize("mode", ["fp64_fused", "mixed_fused"])
def test_tile_orders_and_workgroup_sizes_give_the_same_result(mode):
    """The launch configurations the autotuner chooses among change only how
    tiles meet workgroups and XCDs: contiguous eighths or runs of 1 or 16
    workgroups per XCD, workgroups of 1 or 4 tiles, with graphs.  Every
    slab ends bit for bit the same."""
    cases = small_cases()
    slabs = [c[4] for c in cases]
    cells, box = capacity(slabs)
    w0 = 2 if mode == "fp64_fused" else 3
    out = []
    for block, order in ((64, -1), (64, 0), (256, 0), (256, 4), (128, 4)):
        b = hip.Batch(DT, cells, box, slab_capacity=4, mode=mode)
        ids = [b.add(s) for s in slabs]
        b.set_launch(w0, block, 0, order)
        b.set_graph(10)
        b.run(200)
        out.append([b.vm(i) for i in ids])
        b.close()
    for k in range(1, len(out)):
        for a0, a1 in zip(out[0], out[k]):
            assert np.array_equal(a0, a1)
    say(f"  {mode}: five launch configurations, bit for bit the same")


def test_autotune_restores_state_and_picks_a_candidate():
    cases = small_cases()
    slabs = [c[4] for c in cases]
    cells, box = capacity(slabs)
    res = []
    for tune in (False, True):
        b = hip.Batch(DT, cells, box, slab_capacity=4, mode="mixed_fused", sensors=3)
        ids = [b.add(s) for s in slabs]
        b.run(50)
        if tune:
            best, rows = b.autotune(5)
            say(f"  autotune: best {best:.4f} ms/step, launch {b.launch()}, {len(rows)} candidates")
            assert len(rows) == 18 and best > 0
            assert b.launch() in [tuple(r[:4]) for r in rows]
            assert len({(r[1], r[2], r[3]) for r in rows}) == 18
        b.run(250)
        res.append([np.concatenate([b.vm(i), b.state(i).ravel(), b.t90(i)]) for i in ids])
        assert [b.slab_info(i)["loop"] for i in ids] == [300] * 3
        b.close()
    for k in range(len(slabs)):
        assert np.array_equal(res[0][k], res[1][k])


def test_arena_full_and_reuse():
    cases = small_cases()
    s1, s2 = cases[0][4], cases[1][4]
    c1, x1 = capacity([s1])
    c2, x2 = capacity([s2])
    b = hip.Batch(DT, max(c1, c2), max(x1, x2), slab_capacity=2, mode="mixed_fused")
    i1 = b.add(s1)
    with pytest.raises(hip.ArenaFull):
        b.add(s2)
    b.run(10)
    b.remove(i1)
    i2 = b.add(s2)
    b.run(10)
    assert b.slab_info(i2)["loop"] == 10
    assert b.info()["nslab"] == 1
    b.close()


@pytest.mark.parametrize("mode,parts", [("fp64_fused", 2), ("fp64_fused", 4),
                                        ("mixed_fused", 3)])
def test_split_anatomy_with_halo_exchange_equals_whole(mode, parts):
    a, depth = W.ellipsoid_shell(40, 36, 56, seed=2)
    stims = W.endocardial_point_stimuli(a, depth, 3, seed=2)
    say(f"split along x into {parts} with halo exchange vs whole: {mode}, {a.ncell} cells")
    whole = hip.Slab(a, W.CELLTYPES_EMI, W.DIFFUSION_SCALE, stims, gate_tail_rule=False)
    shares = W.split_x(a, parts)
    slabs = [hip.Slab(s.anatomy, W.CELLTYPES_EMI, W.DIFFUSION_SCALE, stims, gate_tail_rule=False,
                      remote_gid=s.remote_gid, remote_sigma=s.remote_sigma) for s in shares]
    res = []
    for group in ([whole], slabs):
        cells, boxes = capacity(group)
        b = hip.Batch(DT, cells, boxes, slab_capacity=len(group), mode=mode)
        ids = [b.add(s) for s in group]
        linked = 0
        for k in range(1, len(ids)):
            linked += b.link_halo(ids[k], ids[k - 1])
            linked += b.link_halo(ids[k - 1], ids[k])
        if len(ids) > 1:
            assert linked == sum(s.nremote for s in group)
        b.set_graph(10)
        b.run(450)
        vm = np.zeros(a.ncell)
        act = np.zeros(a.ncell)
        st = np.zeros((19, a.ncell))
        gids = [a.gid] if len(ids) == 1 else [s.anatomy.gid for s in shares]
        for sid, g in zip(ids, gids):
            pos = np.searchsorted(a.gid, g)
            vm[pos] = b.vm(sid)
            act[pos] = b.activation(sid)[0]
            st[:, pos] = b.state(sid)
        b.close()
        res.append((vm, act, st))
    (vm0, act0, st0), (vm1, act1, st1) = res
    say(f"  activated {(act0 >= 0).sum()} cells; max |dVm| {np.abs(vm1 - vm0).max():.3e} mV")
    assert (act0 >= 0).sum() > a.ncell // 20
    if mode == "fp64_fused":
        assert np.array_equal(vm1, vm0)
        assert np.array_equal(st1, st0)
        assert np.array_equal(act1, act0)
        # the oracle in the loop: the whole anatomy on the CPU, every cell's
        # gates advanced (a share's reaction groups are not the whole's, so
        # the v2 loop bounds are off on both sides)
        import oracle_lib as OL
        ref = OL.OracleSim(a, W.CELLTYPES_EMI, DT, W.DIFFUSION_SCALE, stims,
                           gate_tail_rule=False)
        ref.run(450)
        want_vm, want_act = ref.vm(), ref.activation()[0]
        ref.close()
        e = np.abs(vm1 - want_vm) / np.maximum(np.abs(want_vm), 1.0)
        say(f"  split run vs oracle: Vm max rel {e.max():.3e}, tact0 identical "
            f"{np.mean(act1 == want_act):.4f}")
        assert e.max() < 1e-9
        assert np.array_equal(act1, want_act)
    else:
        assert np.array_equal(act1 >= 0, act0 >= 0)
        assert np.abs(act1 - act0).max() <= DT * 1.0000001


@pytest.mark.parametrize("mode", ["fp64_fused", "mixed_fused"])
def test_graph_follows_the_live_tiles(mode):
    """The step runs over the live slabs' tiles only: one range from a base,
    or a list when removed slabs leave holes.  A removal and an arrival
    between two runs can leave the same number of live tiles at another
    base, or behind a hole; the replayed graph must then step the new tiles.
    Slabs of 64 tiles each:
      A B C          one range at base 0
      . B C          A leaves: one range at base 64
      D B .          D takes A's place, C leaves: base 0, the same count
      D . E          E takes C's place, B leaves: a list, the same count
    B, D and E end as they do alone."""
    def slab(xmax, ct):
        a = W.brick(16, 16, 16)
        return hip.Slab(a, [ct], W.DIFFUSION_SCALE,
                        [W.box_stimulus(xMax=xmax, nx=16, ny=16, nz=16)])
    sa, sb, sc = slab(2, "endoCellML"), slab(3, "midCellML"), slab(4, "epiCellML")
    sd, se = slab(5, "endoCellML"), slab(6, "midCellML")
    cells, box = capacity([sa, sb, sc])
    say(f"{mode}: graphs across tile moves")
    b = hip.Batch(DT, cells, box, slab_capacity=4, mode=mode)
    b.set_graph(10)
    ia, ib, ic = b.add(sa), b.add(sb), b.add(sc)
    b.run(20)
    b.remove(ia)
    b.run(20)
    idd = b.add(sd)
    b.remove(ic)
    b.run(40)
    ie = b.add(se)
    got = {"b": (b.vm(ib), 80)}
    b.remove(ib)
    b.run(40)
    got["d"] = (b.vm(idd), 80)
    got["e"] = (b.vm(ie), 40)
    b.close()

    def alone(s, n):
        c, x = capacity([s])
        one = hip.Batch(DT, c, x, slab_capacity=1, mode=mode)
        k = one.add(s)
        one.run(n)
        out = one.vm(k)
        one.close()
        return out

    for name, s in (("b", sb), ("d", sd), ("e", se)):
        vm, n = got[name]
        same = np.array_equal(vm, alone(s, n))
        say(f"  slab {name}: {n} steps, identical to alone: {same}")
        assert same


def test_cfg3_sixty_four_slabs_mixed_autotuned_graphs():
    """cfg3 at its own configuration: 64 slabs drawn log-uniform from 32^3
    to 160x160x96 (seed 1; 24.65 M cells), per-slab drug scales on g_Kr,
    g_CaL and g_Na and fibre angles, box stimulus x < 4, mixed precision,
    the launch picked by the autotuner, 10-step graphs, 600 steps.  Eight
    slabs (the smallest, the largest and six between) end bit for bit where
    they end alone; every slab's activation map is sane: times are -1 or in
    (0, 12] ms, a slab that fires fires first in its stimulated planes x < 4
    and its wave moves away from them."""
    nsteps = 600
    cases = W.batched_slabs(64, seed=1)
    slabs = [hip.Slab(c.anatomy, ["endoCellML"], W.DIFFUSION_SCALE, c.stimuli, scales=c.scales)
             for c in cases]
    cells, box = capacity(slabs)
    say(f"cfg3: 64 slabs, {sum(s.ncell for s in slabs)} cells, mixed, autotune, graphs")
    b = hip.Batch(DT, cells, box, slab_capacity=64, mode="mixed_fused")
    ids = [b.add(s) for s in slabs]
    best, table = b.autotune(10)
    say(f"  autotuned {b.launch()} at {best:.3f} ms/step over {len(table)} candidates")
    b.set_graph(10)
    for k in range(3):
        b.run(nsteps // 3)
        b.sync()
        say(f"  {(k + 1) * nsteps // 3} steps")
    order = np.argsort([s.ncell for s in slabs])
    check = sorted(set(order[np.linspace(0, 63, 8).round().astype(int)].tolist()))
    got = {k: (b.vm(ids[k]), b.activation(ids[k])[1]) for k in check}
    acts = [b.activation(ids[k])[1] for k in range(64)]
    b.close()
    end = nsteps * DT
    fired = 0
    for k, (c, act) in enumerate(zip(cases, acts)):
        a = c.anatomy
        x = a.gid % a.nx
        on = act >= 0
        assert np.all((act == -1.0) | ((act > 0) & (act <= end + 1e-9))), k
        if not on.any():
            continue          # drug block (s_Na, s_CaL near 0) can stop a slab firing
        fired += 1
        # the first cells to activate lie in the stimulated planes x < 4, and
        # the wave moves away from them: the mean activation time of the
        # fully activated x-planes does not decrease with x
        assert act[on].min() == act[on & (x < 4)].min(), k
        planes = [act[x == i].mean() for i in range(a.nx) if on[x == i].all()]
        assert np.all(np.diff(planes) >= -DT), k
    say(f"  {fired} of 64 slabs activated")
    assert fired >= 32
    for k in check:
        s = slabs[k]
        c1, x1 = capacity([s])
        one = hip.Batch(DT, c1, x1, slab_capacity=1, mode="mixed_fused")
        sid = one.add(s)
        one.run(nsteps)
        vm, act = one.vm(sid), one.activation(sid)[1]
        one.close()
        same = np.array_equal(vm, got[k][0]) and np.array_equal(act, got[k][1])
        say(f"  slab {k}: {cases[k].anatomy.nx}x{cases[k].anatomy.ny}x{cases[k].anatomy.nz}, "
            f"activated {(act >= 0).mean():.3f}, alone == in the batch: {same}")
        assert same


def test_configuration_A_batch_alone_equals_in_batch_and_other_modes_refuse():
    """Configuration A (the factory's default deck) in the batched engine:
    two slabs with drug scales, each bit for bit what it gives alone after
    400 steps; slabs described with the v2 loop-bound rule run without it,
    as the generic gate loop has none.  The FMA and mixed engines refuse
    the exact functions."""
    cases = W.batched_slabs(n=2, seed=3, lo=(12, 12, 12), hi=(20, 20, 20))
    slabs = [hip.Slab(c.anatomy, ["endoCellML"], W.DIFFUSION_SCALE, c.stimuli, scales=c.scales)
             for c in cases]
    ext = [s.extent() for s in slabs]
    cells, boxes = sum(e[0] for e in ext), sum(e[1] for e in ext)

    def run(which):
        b = hip.Batch(DT, cells, boxes, slab_capacity=2, mode="fp64_fused", exact_functions=True,
                      jhTauSmooth=1)
        ids = [b.add(slabs[k]) for k in which]
        b.run(400)
        out = {k: (b.vm(i), b.state(i), b.activation(i)[1]) for k, i in zip(which, ids)}
        b.close()
        return out

    both = run([0, 1])
    for k in (0, 1):
        alone = run([k])[k]
        say(f"  configuration A slab {k}: {slabs[k].ncell} cells, activated "
            f"{np.mean(alone[2] >= 0):.3f}")
        for x, y in zip(alone, both[k]):
            assert np.array_equal(x, y)
    for mode in ("fp64_fma", "mixed_fused"):
        with pytest.raises(hip.CardioidHipError, match="exact"):
            hip.Batch(DT, cells, boxes, slab_capacity=2, mode=mode, exact_functions=True)
