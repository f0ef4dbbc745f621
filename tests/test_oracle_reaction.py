"""Pins for the oracle's TT06Dev reaction (CPU only).

Three independent anchors:
  * the reference's published single-cell golden (test/results/*/seq/singleCell/
    TT06Opt.stdOut), which exercises the non-gate update and the gate
    update with exact functions (configuration A);
  * the reference's own kernels compiled from source into oracle/_ref
    (configuration B, rational fits + v2 gates), compared bitwise;
  * mathematics: the fits approximate the exact TT06 functions to their
    stated tolerance, gates stay in [0, 1] and relax to their steady state.
"""
import ctypes as C

import numpy as np
import pytest

from cardioid_amd.fitfile import FIT_ORDER, read_fit_file
from oracle_lib import (OraBoxStim, OraCelltype, OraSimParams, aligned, celltype, i64p,
                        make_fitset, oracle, ptr, ref, ref_available, ref_celltype_table,
                        ref_load, u8p)

needs_ref = pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built (reference tree absent)")
FS = read_fit_file()
DT = 0.02

# (mhu_l, mhu_m, tauR_l, tauR_m) hard-coded in each v2 kernel, elec/TT06GatesSimd.c
# (:18-21, :499-502, :806-809, :1130-1133, :1521-1524, :1660-1663, :1973-1976,
#  :2226-2229, :2573-2576, :2874-2877, :3265-3268, :3403-3406, :3722-3725)
V2_ORDERS = [(10, 5, 1, 18), (7, 9, 11, 11), (7, 9, 1, 13), (5, 8, 1, 13), (1, 10, 1, 10),
             (5, 5, 6, 9), (5, 8, 5, 7), (5, 7, 7, 10), (5, 8, 13, 11), (8, 5, 11, 12),
             (6, 12, 1, 1), (7, 8, 6, 7), (7, 8, 11, 11)]


def test_fit_file_matches_v2_kernel_orders():
    assert [f.name for f in FS.fits] == FIT_ORDER
    for eq, (ml, mm, tl, tm) in enumerate(V2_ORDERS):
        fm, ft = FS.fits[7 + 2 * eq], FS.fits[8 + 2 * eq]
        assert (fm.l, fm.m, ft.l, ft.m) == (ml, mm, tl, tm), fm.name


def test_fits_approximate_exact_functions():
    """Each fit stays within a small multiple of its tolerance times the
    function's range on the fit interval: a misplaced coefficient or a
    swapped numerator/denominator fails this by orders of magnitude."""
    O = oracle()
    ofs = make_fitset(FS)
    for i, f in enumerate(FS.fits):
        V = np.arange(f.x0, f.x1 + 1e-9, f.deltaX)
        ex = np.array([O.ora_tt06_exact_func(i, v, 1) for v in V])
        fit = np.array([O.ora_pade_eval(C.byref(ofs.fit[i]), v) for v in V])
        rng = ex.max() - ex.min()
        err = np.abs(fit - ex).max()
        # the fit file was made with mod = 1: its h/j tau fits approximate the
        # Hermite-smoothed rates, which the oracle evaluates at jhTauSmooth = 1
        assert err <= 5 * f.tol * rng, (f.name, err / rng)


def test_singlecell_golden_configuration_A():
    """test/tests/singleCell: singleCell -o TT06Opt.data -h 0.01 -N 4, epiCellML,
    exact functions; stimulus +60 for 1 ms at t = 0; Vm every 1 ms.  The
    reference compares with numCompare at 1e-7 (range-normalised)."""
    O = oracle()
    gold = np.loadtxt("tests/golden/TT06Opt_singleCell_epi_dt0.01.stdOut", skiprows=1)
    ct = celltype("epiCellML")
    cts = (OraCelltype * 1)(ct)
    gid = np.zeros(1, np.int64)
    ctype = np.zeros(1, np.uint8)
    sig = np.zeros(6)
    ofs = make_fitset(FS)
    p = OraSimParams(nx=1, ny=1, nz=1, dx=0.1, dy=0.1, dz=0.1, dt=0.01, diffusionScale=1.0,
                     ncell=1, gid=ptr(gid, i64p), celltype=ptr(ctype, u8p), sigma=ptr(sig),
                     ncelltype=1, celltypes=cts, scales=None, fits=C.pointer(ofs), exact=1,
                     threshold=-40.0, nthreads=1)
    st = OraBoxStim(xMin=-1, xMax=1, yMin=-1, yMax=1, zMin=-1, zMax=1, t0=-1000, tf=1e30,
                    period=1000, duration=1, amplitude=60.0, phase=0.0)
    s = O.ora_sim_create(C.byref(p), 1, C.byref(st))
    vm = np.zeros(1)
    out = []
    for k in range(gold.shape[0]):
        O.ora_sim_get_vm(s, ptr(vm))
        out.append(vm[0])
        if k + 1 < gold.shape[0]:
            O.ora_sim_run(s, 100)
    O.ora_sim_destroy(s)
    out = np.array(out)
    rng = gold[:, 1].max() - gold[:, 1].min()
    assert np.abs(out - gold[:, 1]).max() / rng < 1e-7
    # in fact it agrees to the printed digits
    assert np.abs(out - gold[:, 1]).max() < 1e-12


def _random_cells(ct, n, P, rng):
    O = oracle()
    st = aligned(19 * P)
    Vm = aligned(n)
    O.ora_tt06_init_state(C.byref(ct), n, ptr(st), P, ptr(Vm))
    Vm[:] = rng.uniform(-90, 40, n)
    S = st.reshape(19, P)[:, :n]
    S[:] *= rng.uniform(0.8, 1.2, (19, n))
    S[1] = 138.0 * rng.uniform(0.95, 1.05, n) / O.ora_c9() + Vm
    return st, Vm


@needs_ref
@pytest.mark.parametrize("config", ["B", "A", "A-mod"])
@pytest.mark.parametrize("name", ["endoCellML", "midCellML", "epiCellML", "midRRG"])
def test_reaction_bitwise_against_reference_kernels(config, name):
    """Configuration B (fits, v2 gates) and configuration A (exact
    functions, generic gates) with mod = 0 and mod = 1 (smoothed h and j
    rates), each against the reference's own kernels."""
    exact = config != "B"
    mod = 0 if config == "A" else 1
    O = oracle()
    R = ref_load(FS, DT, exact=exact, jhTauSmooth=mod)
    ofs = make_fitset(FS, DT)
    ct = celltype(name)
    R.ref_set_celltype(ct.s_switch, ct.P_NaK, ct.g_Ks, ct.g_Kr, ct.g_to, ct.g_NaL)
    rng = np.random.default_rng(11)
    n, P = 1001, 1004
    st, Vm = _random_cells(ct, n, P, rng)
    st2 = aligned(19 * P)
    st2[:] = st
    d1, d2 = aligned(n), aligned(n)
    R.ref_nongate(DT, n, P, ptr(Vm), ptr(st), ptr(d1))
    if exact:
        O.ora_tt06_nongate_exact(DT, C.byref(ct), None, n, ptr(Vm), ptr(st2), P, ptr(d2))
    else:
        O.ora_tt06_nongate(C.byref(ofs), DT, C.byref(ct), None, n, ptr(Vm), ptr(st2), P, ptr(d2))
    assert np.array_equal(d1, d2)
    assert np.array_equal(st, st2)
    if exact:
        R.ref_gates_generic(DT, n, P, ptr(Vm), ptr(st))
        O.ora_tt06_gates_exact(DT, ct.s_switch, mod, n, ptr(Vm), ptr(st2), P)
    else:
        R.ref_gates_fast(DT, n, P, ptr(Vm), ptr(st))
        O.ora_tt06_gates_v2(C.byref(ofs), ct.s_switch, n, ptr(Vm), ptr(st2), P)
    assert np.array_equal(st, st2)


@needs_ref
@pytest.mark.parametrize("mod", [0, 1])
def test_exact_functions_bitwise_against_reference(mod):
    """All 33 exact functions of configuration A against the reference's
    (TT06Func.cc:261-417, TT06NonGates.c:319-364, TT06Tau.cc), over
    V in [-110, 140] on a 0.01 mV grid plus the ends of the smoothed h and
    j intervals and their neighbouring doubles."""
    O, R = oracle(), ref()
    V = list(np.arange(-110.0, 140.0, 0.01))
    for x in (-64.20, -23.3, -48.85, -17.6, -40.0):
        V += [np.nextafter(x, -np.inf), x, np.nextafter(x, np.inf)]
    for i in range(33):
        a = np.array([O.ora_tt06_exact_func(i, v, mod) for v in V])
        b = np.array([R.ref_exact_func(i, v, mod) for v in V])
        assert np.array_equal(a, b), (i, mod)
    # the smoothing cubic meets the raw rate at both ends of its interval
    c = np.zeros(6)
    for which, idx in ((0, 10), (1, 12)):
        O.ora_tt06_mod_coefficients(which, ptr(c))
        x0, x1 = c[0], c[1]
        for x in (x0, x1):
            raw = O.ora_tt06_exact_func(idx, np.nextafter(x, -np.inf if x == x0 else np.inf), 0)
            assert abs(O.ora_tt06_exact_func(idx, x, 1) - raw) <= 1e-9 * abs(raw)


@needs_ref
def test_celltype_tables_equal_the_reference():
    """Every row of TT06Func::getStandardCellTypes (TT06Func.cc:421-612), as
    the reference build returns it, equals the oracle's table and the HIP
    library's (cdh_celltype_standard) bit for bit: s_switch, the five
    conductances, the rest Vm and the 19 rest states."""
    from cardioid_amd import hip
    table = ref_celltype_table()
    assert set(table) == {"endoCellML", "midCellML", "epiCellML", "endoRRG", "midRRG", "epiRRG"}
    for name, (sw, parms, state) in table.items():
        for ct in (celltype(name), hip.standard_celltype(name)):
            got = np.array([ct.P_NaK, ct.g_Ks, ct.g_Kr, ct.g_to, ct.g_NaL, ct.Vm0])
            assert ct.s_switch == sw, name
            assert np.array_equal(got, parms), (name, got, parms)
            assert np.array_equal(np.array(ct.state0[:]), state), name


def test_threaded_oracle_equals_its_serial_run():
    """The reference advances each reaction group on one thread; the oracle
    splits a group across threads in chunks that start at multiples of 12,
    where the v2 loops' per-group rules (tail bound, f-gate form) fall as in
    the whole group.  1, 5 and 8 threads give the same doubles, with and
    without the loop bounds."""
    from cardioid_amd import workloads as W
    from oracle_lib import OracleSim
    a, depth = W.ellipsoid_shell(24, 22, 30, seed=3)
    stims = W.endocardial_point_stimuli(a, depth, 3, seed=3)
    for rule in (True, False):
        out = []
        for nt in (1, 5, 8):
            r = OracleSim(a, ["endoCellML", "midCellML", "epiCellML"], DT, 714.2857143, stims,
                          nthreads=nt, gate_tail_rule=rule)
            r.run(120)
            out.append((r.vm(), r.state()))
            r.close()
        for vm, st in out[1:]:
            assert np.array_equal(vm, out[0][0]) and np.array_equal(st, out[0][1])


@needs_ref
@pytest.mark.parametrize("threads", [(1, 1), (6, 6)])
def test_loop_with_the_reference_reaction_equals_the_oracle(threads):
    """The omp loop with the reaction run by the reference's own kernels
    (one thread per group as the reference runs it, or split in 12-cell
    multiples) equals the oracle's loop bit for bit: Vm, all 19 states and
    the activation times, over three cell types."""
    from cardioid_amd import workloads as W
    from oracle_lib import OracleSim
    a, depth = W.ellipsoid_shell(20, 18, 26, seed=5)
    stims = W.endocardial_point_stimuli(a, depth, 2, seed=5)
    cts = ["endoCellML", "midCellML", "epiCellML"]
    own = OracleSim(a, cts, DT, 714.2857143, stims, nthreads=6)
    refk = OracleSim(a, cts, DT, 714.2857143, stims, nthreads=6)
    refk.use_reference_reaction(*threads)
    own.run(300)
    refk.run(300)
    assert np.array_equal(own.vm(), refk.vm())
    assert np.array_equal(own.state(), refk.state())
    for x, y in zip(own.activation(), refk.activation()):
        assert np.array_equal(x, y)
    assert (own.activation()[1] >= 0).any()


@needs_ref
def test_each_v2_gate_bitwise_over_group_sizes():
    """Group sizes cover every split of the 8/12/16-wide main loops, the
    4-wide tail loop and the unvisited remainder."""
    O = oracle()
    R = ref_load(FS, DT)
    ofs = make_fitset(FS, DT)
    rng = np.random.default_rng(5)
    sizes = list(range(1, 50)) + [1003, 1004, 1005, 1006, 1009, 1013, 1017, 1020, 1024, 4103]
    for n in sizes:
        for eq in range(13):
            Vm = aligned(n)
            Vm[:] = rng.uniform(-95, 50, n)
            g = aligned(n)
            g[:] = rng.uniform(0, 1, n)
            g2 = aligned(n)
            g2[:] = g
            R.ref_gate_v2(DT, eq, n, ptr(Vm), ptr(g))
            O.ora_tt06_gate_v2(C.byref(ofs), eq, n, ptr(Vm), ptr(g2))
            assert np.array_equal(g, g2), (n, eq)


@needs_ref
def test_fit_evaluation_bitwise():
    O = oracle()
    R = ref_load(FS, DT)
    ofs = make_fitset(FS, DT)
    V = np.linspace(-110, 140, 2501)
    for i in range(33):
        a = np.array([R.ref_fit_eval(i, v) for v in V])
        b = np.array([O.ora_pade_eval(C.byref(ofs.fit[i]), v) for v in V])
        assert np.array_equal(a, b), FS.fits[i].name


def test_gates_bounded_and_relax_to_steady_state():
    O = oracle()
    ofs = make_fitset(FS, DT)
    rng = np.random.default_rng(3)
    n = 4000
    for eq in range(13):
        Vm = aligned(n)
        Vm[:] = rng.uniform(-90, 40, n)
        g = aligned(n)
        g[:] = rng.uniform(0, 1, n)
        O.ora_tt06_gate_v2(C.byref(ofs), eq, n, ptr(Vm), ptr(g))
        assert np.all(g >= -1e-12) and np.all(g <= 1 + 1e-9), eq
    # clamped voltage: the distance to mhu(V) contracts geometrically,
    # e_K = (1 - r)^K e_0 with r = dt/tau (r = 1 - exp(-dt/tau) for m)
    n = 64
    K = 200
    Vm = aligned(n)
    Vm[:] = np.linspace(-85, 20, n)
    for eq in range(13):
        g = aligned(n)
        g[:] = 0.5
        for _ in range(K):
            O.ora_tt06_gate_v2(C.byref(ofs), eq, n, ptr(Vm), ptr(g))
        mhu = np.array([O.ora_pade_eval(C.byref(ofs.fit[7 + 2 * eq]), v) for v in Vm])
        x = np.array([O.ora_pade_eval(C.byref(ofs.fit[8 + 2 * eq]), v) for v in Vm])
        r = 1 - np.exp(-x) if eq == 0 else x
        expect = mhu + (0.5 - mhu) * (1 - r) ** K
        assert np.abs(g - expect).max() < 1e-9, eq


def test_m_gate_rush_larsen_matches_exponential():
    """The m gate uses Rush-Larsen with exp(x) replaced by its degree-16
    Taylor polynomial (elec/TT06GatesSimd.c:10-14, :239-241): the update
    equals mhu*(1 - 1/T16(x)), and matches 1 - exp(-x) where the series
    converges (x = dt/tau reaches ~23 at dt = 0.02, where 1/T16 ~ 0)."""
    import math
    O = oracle()
    ofs = make_fitset(FS, DT)
    n = 200
    Vm = aligned(n)
    Vm[:] = np.linspace(-100, 50, n)
    g = aligned(n)
    g[:] = 0.0
    O.ora_tt06_gate_v2(C.byref(ofs), 0, n, ptr(Vm), ptr(g))
    mhu = np.array([O.ora_pade_eval(C.byref(ofs.fit[7]), v) for v in Vm])
    x = np.array([O.ora_pade_eval(C.byref(ofs.fit[8]), v) for v in Vm])
    t16 = sum(x ** k / math.factorial(k) for k in range(17))
    assert np.abs(g - mhu * (1 - 1 / t16)).max() < 1e-14
    small = x < 2.0
    assert small.sum() > 20
    assert np.abs(g - mhu * (1 - np.exp(-x)))[small].max() < 1e-9


def test_gate_flags_rule():
    O = oracle()
    f = [O.ora_gate_flags(i, 27) for i in range(27)]
    # v2 loops stop at 4*floor(27/4) = 24; the f-gate main loop covers 0..23
    assert [x & 1 for x in f] == [0] * 24 + [1] * 3
    assert [(x >> 1) & 1 for x in f] == ([1] * 4 + [0] * 8) * 2 + [0] * 3


def test_flop_count_rederived_from_the_fit_orders():
    """1,060 flop per endocardial cell-step: 66 Horner multiply-adds in the
    non-gate fits and 303 in the 24 gate fits of the fixture, plus the fixed
    arithmetic; the s1 fits of mid/epi cells add 9 multiply-adds."""
    from cardioid_amd import flops
    b0, b1 = flops.breakdown(s_switch=0), flops.breakdown(s_switch=1)
    assert (b0["nongate_fit_fmas"], b0["gate_fit_fmas"]) == (66, 303)
    assert b0["nongate"] == 309 and b0["gates"] == 706 and b0["reaction"] == 1019
    assert b0["total"] == flops.flop_per_cell_step() == 1060
    assert b1["gate_fit_fmas"] - b0["gate_fit_fmas"] == 9
