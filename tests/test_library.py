"""CPU checks of the HIP library: it loads, exports every entry point the C
header declares, and rejects bad input on the host side before touching a
device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from cardioid_amd import hip
from cardioid_amd.fitfile import read_fit_file

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cardioid_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cdh_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = hip.lib()
    names = declared_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(hip.EXPORTS)


def test_standard_celltypes_and_checkpoint_fields():
    ct = hip.standard_celltype("endoCellML")
    assert ct.s_switch == 0 and abs(ct.Vm0 + 86.709) < 1e-12
    assert hip.standard_celltype("epiCellML").s_switch == 1
    with pytest.raises(hip.CardioidHipError):
        hip.standard_celltype("noSuchCell")
    L = hip.lib()
    names = []
    for i in range(19):
        n, u = C.c_char_p(), C.c_char_p()
        h = C.c_int()
        assert L.cdh_tt06_checkpoint_field(i, C.byref(n), C.byref(u), C.byref(h)) == 0
        names.append(n.value.decode())
        assert L.cdh_tt06_var_handle(n.value) == h.value == i
    # the 19 names of TT06Dev_Reaction::getHandleMap (elec/TT06Dev_Reaction.cc:533-551)
    assert set(names) == {"dVK_i", "Na_i", "Ca_i", "Xr1_gate", "Xr2_gate", "Xs_gate", "m_gate",
                          "h_gate", "j_gate", "Ca_ss", "d_gate", "f_gate", "f2_gate",
                          "fCass_gate", "s_gate", "r_gate", "Ca_SR", "R_prime", "jL_gate"}
    assert L.cdh_tt06_var_handle(b"bogus") == -1


def _tt06_params(**kw):
    fs = read_fit_file()
    l, m, c = fs.packed()
    p = hip.TT06Params(dt=0.02, tolerance=1e-4, mod=1, fastGate=1, fastNonGate=0,
                       fit_l=hip.ptr(l, hip.ip), fit_m=hip.ptr(m, hip.ip), fit_coef=hip.ptr(c),
                       celltype=hip.standard_celltype("endoCellML"), s_kr=1.0, s_cal=1.0, s_na=1.0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p, (l, m, c)


@pytest.mark.parametrize("field,value", [("tolerance", 1e-3), ("mod", 0), ("fastGate", 0),
                                         ("fastNonGate", 1)])
def test_reaction_rejects_unsupported_configurations(field, value):
    """With rational fits only configuration B is taken; tolerance 0
    (configuration A, the exact functions) is accepted and needs a device."""
    p, keep = _tt06_params(**{field: value})
    h = C.c_void_p()
    rc = hip.lib().cdh_tt06_create(C.byref(p), 16, C.byref(h))
    assert rc == 3, hip.lib().cdh_last_error()
    assert b"TT06Dev" in hip.lib().cdh_last_error()


def test_reaction_rejects_fit_orders_the_kernels_do_not_have():
    fs = read_fit_file()
    fs.fits[8].l += 1     # mTauR
    fs.fits[8].coef = np.append(fs.fits[8].coef, 0.0)
    l, m, c = fs.packed()
    p, _ = _tt06_params()
    p.fit_l, p.fit_m, p.fit_coef = hip.ptr(l, hip.ip), hip.ptr(m, hip.ip), hip.ptr(c)
    h = C.c_void_p()
    assert hip.lib().cdh_tt06_create(C.byref(p), 16, C.byref(h)) == 3
    assert b"orders" in hip.lib().cdh_last_error()


def test_null_arguments_are_errors_not_crashes():
    L = hip.lib()
    assert L.cdh_tt06_calc(None, 0.02, None, 0, None, None, None, None) == 1
    assert L.cdh_fgr_calc(None, None, None) == 1
    assert L.cdh_sim_run(None, 1) == 1
    assert L.cdh_fill(-1, 0.0, None, None) == 1


def _slab(gid, n=(4, 4, 4), **kw):
    from cardioid_amd import workloads as W
    gid = np.asarray(gid, dtype=np.int64)
    a = W.Anatomy(n[0], n[1], n[2], 0.2, 0.2, 0.2, gid, np.zeros(gid.size, np.uint8),
                  np.tile([1e-4, 0, 0, 1e-4, 0, 1e-4], (gid.size, 1)))
    return hip.Slab(a, ["endoCellML"], 714.0, **kw)


def test_slab_limits_are_host_errors():
    """Slab descriptions the engine cannot hold fail on the host, before any
    device work: a padded box of 2^28 points or more, a box plane of 2^22
    points or more, cells outside the grid, bad cell-type indices, more than 8
    stimuli, remote cells outside the grid."""
    from cardioid_amd import workloads as W
    ok = _slab([0, 1, 2])
    assert ok.extent()[:2] == (64, 5 * 3 * 3)
    n = (700, 700, 700)
    with pytest.raises(hip.CardioidHipError, match="2\\^28"):
        _slab([0, 700 ** 3 - 1], n=n).extent()
    n = (1, 2098, 2098)       # box 3 x 2100 x 2100: a y-z plane of 4.41 M points
    with pytest.raises(hip.CardioidHipError, match="2\\^22"):
        _slab([0, 2098 * 2098 - 1], n=n).extent()
    with pytest.raises(hip.CardioidHipError, match="outside the grid"):
        _slab([0, 64]).extent()
    bad = _slab([0, 1])
    bad._keep[2][1] = 3       # cell-type index beyond the one type given
    with pytest.raises(hip.CardioidHipError, match="cell type index"):
        bad.extent()
    stims = [W.box_stimulus(nx=4, ny=4, nz=4)] * 9
    with pytest.raises(hip.CardioidHipError, match="8 stimuli"):
        _slab([0, 1], stimuli=stims).extent()
    with pytest.raises(hip.CardioidHipError, match="remote gid"):
        _slab([0, 1], remote_gid=[64], remote_sigma=np.zeros((1, 6))).extent()


def test_checkpoint_of_no_cells(tmp_path):
    path = hip.write_checkpoint(str(tmp_path), 3, 0.06, (4, 4, 4), np.zeros(0, np.int64),
                                np.zeros(0), np.zeros((19, 0)))
    vm, st, loop, time = hip.read_checkpoint(path, np.zeros(0, np.int64))
    assert vm.size == 0 and st.shape == (19, 0) and (loop, time) == (3, 0.06)


def test_stream_probe_rejects_bad_arguments():
    g = C.c_double()
    L = hip.lib()
    assert L.cdh_stream_rate(3, 1024, 1, C.byref(g)) == 1
    assert L.cdh_stream_rate(0, 0, 1, C.byref(g)) == 1
    assert L.cdh_stream_rate(2, 1024, 0, C.byref(g)) == 1
    assert L.cdh_stream_rate(1, 1024, 1, None) == 1
    with pytest.raises(KeyError):
        hip.stream_rate("copy", 1024)
