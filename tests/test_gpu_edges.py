"""Edge cases of the drop-in objects and the engine on an MI355X: objects of
zero cells (a rank with no cells of a type, elec/ReactionManager.cc:184-198),
a diffusion object whose only neighbours are remote cells, and a reaction
object that is driven through its separate non-gate and gate updates (the
pdr loop's entry points) and lands where calc lands.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from cardioid_amd import hip
from cardioid_amd.fitfile import read_fit_file

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


def vp(t):
    return C.c_void_p(t.data_ptr())


def _tt06(n, fs=read_fit_file()):
    l, m, c = fs.packed()
    p = hip.TT06Params(dt=0.02, tolerance=1e-4, mod=1, fastGate=1, fastNonGate=0,
                       fit_l=hip.ptr(l, hip.ip), fit_m=hip.ptr(m, hip.ip), fit_coef=hip.ptr(c),
                       celltype=hip.standard_celltype("epiCellML"), s_kr=1.0, s_cal=1.0,
                       s_na=1.0)
    h = C.c_void_p()
    hip.check(hip.lib().cdh_tt06_create(C.byref(p), n, C.byref(h)), "cdh_tt06_create")
    return h, (l, m, c, p)


def test_zero_cell_objects():
    say("drop-in objects with no cells")
    L = hip.lib()
    h, keep = _tt06(0)
    vm = torch.zeros(1, dtype=torch.float64, device="cuda")
    hip.check(L.cdh_tt06_calc(h, 0.02, None, 0, vp(vm), vp(vm), vp(vm), None), "calc")
    hip.check(L.cdh_tt06_update_gate(h, 0.02, None, 0, vp(vm), None), "gate")
    hip.check(L.cdh_tt06_get_state(h, hip.ptr(np.zeros(1))), "get_state")
    L.cdh_tt06_destroy(h)
    gid = np.zeros(0, np.int64)
    ct = np.zeros(0, np.int32)
    sig = np.zeros(0)
    p = hip.FGRParams(nx=4, ny=4, nz=4, dx=0.2, dy=0.2, dz=0.2, diffusionScale=714.0, nLocal=0,
                      nRemote=0, gid=hip.ptr(gid, hip.i64p), cellType=ct.ctypes.data_as(hip.ip),
                      sigma=hip.ptr(sig))
    d = C.c_void_p()
    hip.check(L.cdh_fgr_create(C.byref(p), C.byref(d)), "cdh_fgr_create")
    hip.check(L.cdh_fgr_update_local_voltage(d, vp(vm), None), "update")
    hip.check(L.cdh_fgr_calc(d, vp(vm), None), "calc")
    L.cdh_fgr_destroy(d)
    s = C.c_void_p()
    hip.check(L.cdh_activation_sensor_create(0, -40.0, C.byref(s)), "sensor")
    hip.check(L.cdh_activation_sensor_eval(s, 0.02, vp(vm), None), "eval")
    L.cdh_activation_sensor_destroy(s)
    torch.cuda.synchronize()


def test_split_reaction_updates_equal_calc():
    """updateNonGate then updateGate (the pdr loop, elec/TT06Dev_Reaction.cc:
    408-453) give the states and dVm that calc gives."""
    say("nongate + gate updates vs calc, 777 cells")
    L = hip.lib()
    n = 777
    rng = np.random.default_rng(3)
    vm = rng.uniform(-90, 30, n)
    out = []
    for split in (False, True):
        h, keep = _tt06(n)
        dvm = torch.zeros(n, dtype=torch.float64, device="cuda")
        dv = torch.from_numpy(vm).cuda()
        for _ in range(20):
            if split:
                hip.check(L.cdh_tt06_update_nongate(h, 0.02, None, n, vp(dv), vp(dvm), None), "ng")
                hip.check(L.cdh_tt06_update_gate(h, 0.02, None, n, vp(dv), None), "g")
            else:
                hip.check(L.cdh_tt06_calc(h, 0.02, None, n, vp(dv), vp(dv), vp(dvm), None), "c")
        torch.cuda.synchronize()
        st = np.zeros(19 * n)
        hip.check(L.cdh_tt06_get_state(h, hip.ptr(st)), "state")
        out.append((st, dvm.cpu().numpy()))
        L.cdh_tt06_destroy(h)
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def test_remote_only_neighbours_and_value_accessors():
    """A diffusion object of one local cell whose six face neighbours are all
    remote: dVm comes from the remote voltages alone (closed form on a uniform
    isotropic conductivity: sigma/h^2 * sum(phi_nbr - phi_0) * scale), and
    the reaction's getValue / setValue address the same state the bulk
    transfer does."""
    L = hip.lib()
    nx = ny = nz = 3
    centre = 1 + 3 * (1 + 3 * 1)
    nbr = [centre - 1, centre + 1, centre - 3, centre + 3, centre - 9, centre + 9]
    gid = np.array([centre] + nbr, np.int64)
    ct = np.ones(7, np.int32)
    s = 1e-4
    sig = np.tile([s, 0, 0, s, 0, s], 7).astype(np.float64)
    p = hip.FGRParams(nx=nx, ny=ny, nz=nz, dx=0.2, dy=0.2, dz=0.2, diffusionScale=714.0,
                      nLocal=1, nRemote=6, gid=hip.ptr(gid, hip.i64p),
                      cellType=ct.ctypes.data_as(hip.ip), sigma=hip.ptr(sig))
    d = C.c_void_p()
    hip.check(L.cdh_fgr_create(C.byref(p), C.byref(d)), "cdh_fgr_create")
    local = torch.tensor([-80.0], dtype=torch.float64, device="cuda")
    remote = torch.tensor([-70.0, -60.0, -50.0, -40.0, -30.0, -20.0], dtype=torch.float64,
                          device="cuda")
    dvm = torch.zeros(1, dtype=torch.float64, device="cuda")
    hip.check(L.cdh_fgr_update_local_voltage(d, vp(local), None), "local")
    hip.check(L.cdh_fgr_update_remote_voltage(d, vp(remote), None), "remote")
    hip.check(L.cdh_fgr_calc(d, vp(dvm), None), "calc")
    torch.cuda.synchronize()
    want = s / 0.2 ** 2 * float((remote + 80.0).sum()) * 714.0
    got = float(dvm.cpu()[0])
    say(f"  dVm {got:.6e}, closed form {want:.6e}")
    # the weights are floats: one rounding of sigma/h^2 each (2^-24 relative)
    assert abs(got - want) <= 2e-7 * abs(want)
    L.cdh_fgr_destroy(d)
    h, keep = _tt06(5)
    st = np.zeros(19 * 5)
    hip.check(L.cdh_tt06_get_state(h, hip.ptr(st)), "state")
    v = C.c_double()
    for cell, handle in ((0, 0), (4, 18), (2, 7)):
        hip.check(L.cdh_tt06_get_value(h, cell, handle, C.byref(v)), "get")
        assert v.value == st[handle * 5 + cell]
        hip.check(L.cdh_tt06_set_value(h, cell, handle, 0.125 + cell + handle), "set")
    hip.check(L.cdh_tt06_get_state(h, hip.ptr(st)), "state")
    for cell, handle in ((0, 0), (4, 18), (2, 7)):
        assert st[handle * 5 + cell] == 0.125 + cell + handle
    L.cdh_tt06_destroy(h)


def test_scatter_and_sensor_take_unaligned_arrays():
    """The diffusion object's voltage scatter and the activation sensor read
    four cells per thread with 16-byte loads when the caller's array allows
    it, and one cell per thread when it does not (an array starting 8 bytes
    into an allocation).  Both give the same bits, over a ragged cell count
    (4199 = 4 * 1049 + 3)."""
    L = hip.lib()
    nx, ny, nz = 19, 17, 13
    n = nx * ny * nz
    say(f"scatter and sensor, aligned and unaligned arrays, {n} cells")
    gid = np.arange(n, dtype=np.int64)
    ct = np.ones(n, np.int32)
    sig = np.tile([1.3e-4, 0.2e-4, 0.1e-4, 1.1e-4, 0.3e-4, 0.9e-4], n).astype(np.float64)
    p = hip.FGRParams(nx=nx, ny=ny, nz=nz, dx=0.2, dy=0.2, dz=0.2, diffusionScale=714.0,
                      nLocal=n, nRemote=0, gid=hip.ptr(gid, hip.i64p),
                      cellType=ct.ctypes.data_as(hip.ip), sigma=hip.ptr(sig))
    rng = np.random.default_rng(11)
    steps = [rng.uniform(-90, 40, n) for _ in range(6)]
    out = []
    for shift in (0, 1):
        d = C.c_void_p()
        hip.check(L.cdh_fgr_create(C.byref(p), C.byref(d)), "cdh_fgr_create")
        s = C.c_void_p()
        hip.check(L.cdh_activation_sensor_create(n, -40.0, C.byref(s)), "sensor")
        buf = torch.zeros(n + 2, dtype=torch.float64, device="cuda")
        vm = buf[shift:shift + n]
        assert (vm.data_ptr() % 16 == 0) == (shift == 0)
        dvms = []
        for k, v in enumerate(steps):
            vm.copy_(torch.from_numpy(v))
            dvm = torch.zeros(n, dtype=torch.float64, device="cuda")
            hip.check(L.cdh_fgr_update_local_voltage(d, vp(vm), None), "update")
            hip.check(L.cdh_fgr_calc(d, vp(dvm), None), "calc")
            hip.check(L.cdh_activation_sensor_eval(s, 0.02 * (k + 1), vp(vm), None), "eval")
            dvms.append(dvm.cpu().numpy())
        torch.cuda.synchronize()
        act = [np.zeros(n) for _ in range(3)]
        hip.check(L.cdh_activation_sensor_get(s, *(hip.ptr(a) for a in act)), "get")
        out.append((dvms, act))
        L.cdh_activation_sensor_destroy(s)
        L.cdh_fgr_destroy(d)
    for a, b in zip(out[0][0], out[1][0]):
        assert np.array_equal(a, b)
    for a, b in zip(out[0][1], out[1][1]):
        assert np.array_equal(a, b)
    # the sensors saw crossings: some times recorded, not all
    assert 0 < np.count_nonzero(out[0][1][1] >= 0) < n
