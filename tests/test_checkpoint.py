"""Checkpoint / restart files (elec/checkpointIO.cc) through the C ABI, on the
CPU: the writer and reader are host code.

Pinned by: the reference's own pio state file (a slice of
test/tests/diffusionTest/snapshot.000000000001/state#000000, whose values
follow test/tests/diffusionTest/initialConditions.py), the field list and
units of TT06Dev_Reaction.cc:528-552 in std::map order
(ReactionManager.cc:384-399), and the record formats of
checkpointIO.cc:137-189.
"""
import os
import re
import struct

import numpy as np
import pytest

from cardioid_amd import hip

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "pio_diffusion", "state#000000")

# TT06Dev_Reaction.cc:533-551: name -> unit
REF_UNITS = {
    "dVK_i": "mV", "Na_i": "mM", "Ca_i": "mM", "Xr1_gate": "1", "Xr2_gate": "1",
    "Xs_gate": "1", "m_gate": "1", "h_gate": "1", "j_gate": "1", "Ca_ss": "mM",
    "d_gate": "1", "f_gate": "1", "f2_gate": "1", "fCass_gate": "1", "s_gate": "1",
    "r_gate": "1", "Ca_SR": "mM", "R_prime": "1", "jL_gate": "mM",
}


def _lib_or_skip():
    try:
        return hip.lib()
    except hip.CardioidHipError as e:
        pytest.skip(str(e))


def _slot_names():
    L = _lib_or_skip()
    import ctypes as C
    out = []
    for i in range(hip.NSTATE):
        nm, un, h = C.c_char_p(), C.c_char_p(), C.c_int()
        assert L.cdh_tt06_checkpoint_field(i, C.byref(nm), C.byref(un), C.byref(h)) == 0
        out.append((nm.value.decode(), un.value.decode()))
    return out


def _random_case(n, seed=0):
    rng = np.random.default_rng(seed)
    gid = rng.permutation(40 * 40 * 40)[:n].astype(np.int64)
    vm = rng.uniform(-90, 40, n)
    st = rng.uniform(0, 1, (hip.NSTATE, n)) * 10.0 ** rng.integers(-6, 3, (hip.NSTATE, 1))
    return gid, vm, st


def _header(path):
    raw = open(path, "rb").read()
    end = raw.index(b"}\n\n") + 3
    text = raw[:end].decode()
    kv = {}
    body = text[text.index("{") + 1:text.rindex("}")]
    for item in body.split(";"):
        if "=" in item:
            k, v = item.split("=", 1)
            kv[k.strip()] = v.strip()
    return text, kv, raw[end:]


def test_reference_state_file():
    """The reference's own ASCII state file (gid Vm, lrec 35) reads with the
    values its generator gives every gid, in any cell order."""
    _lib_or_skip()
    text, kv, body = _header(FIXTURE)
    n = int(kv["nrecord"])
    gids = np.array([int(l.split()[0]) for l in body.decode().splitlines()], dtype=np.int64)
    assert gids.size == n
    order = np.random.default_rng(3).permutation(n)
    cells = gids[order]
    vm0 = np.full(n, -123.0)
    st0 = np.random.default_rng(4).uniform(size=(hip.NSTATE, n))
    import ctypes as C
    vm, st = vm0.copy(), st0.copy()
    loop, time = C.c_int(), C.c_double()
    hip.check(hip.lib().cdh_checkpoint_read(FIXTURE.encode(), n, hip.ptr(cells, hip.i64p),
                                            hip.ptr(vm), hip.ptr(st), C.byref(loop),
                                            C.byref(time)), "read")
    nx = ny = 40
    expect = (cells % nx + cells // nx % ny + cells // (nx * ny)) % 3
    np.testing.assert_array_equal(vm, expect.astype(float))
    # the file has no reaction fields: states keep their values
    np.testing.assert_array_equal(st, st0)
    assert (loop.value, time.value) == (1, 1.0)
    # the same through "<base>#" and "<base>"
    base = FIXTURE[:-len("#000000")]
    for name in (base + "#", base):
        vm2, _, l2, t2 = hip.read_checkpoint(name, cells)
        np.testing.assert_array_equal(vm2, vm)


def test_field_list_matches_reference():
    names = _slot_names()
    assert dict(names) == REF_UNITS


@pytest.mark.parametrize("ascii", [True, False])
def test_writer_format(tmp_path, ascii):
    _lib_or_skip()
    n = 37
    gid, vm, st = _random_case(n)
    path = hip.write_checkpoint(str(tmp_path), 1234, 12.34, (40, 40, 40), gid, vm, st, ascii=ascii,
                                simulate_name="simulate")
    snap = tmp_path / "snapshot.000000001234"
    assert path == str(snap / "state")
    text, kv, body = _header(path + "#000000")
    fields = sorted(REF_UNITS)          # std::map order
    assert kv["datatype"] == ("FIXRECORDASCII" if ascii else "FIXRECORDBINARY")
    assert kv["nfiles"] == "1" and kv["nrecord"] == str(n)
    assert kv["endian_key"] == str(struct.unpack("<i", b"1234")[0])
    assert kv["time"] == "12.340000" and kv["loop"] == "1234"
    assert kv["reactionMethod"] == "TT06Dev" and kv["nfields"] == str(len(fields) + 2)
    assert kv["field_names"] == "gid Vm " + " ".join(fields)
    ftype = "f" if ascii else "f8"
    assert kv["field_types"] == ("u f " if ascii else "u8 f8 ") + " ".join([ftype] * len(fields))
    assert kv["field_units"] == "1 mV " + " ".join(REF_UNITS[f] for f in fields)
    assert re.search(r"\n   nx = 40; ny = 40; nz = 40;\n}\n\n$", text)
    assert text.startswith("state FILEHEADER {\n   create_time = ")
    slot = {nm: i for i, (nm, _) in enumerate(_slot_names())}
    lrec = int(kv["lrec"])
    if ascii:
        assert lrec == 34 + 22 * len(fields) + 1
        assert len(body) == n * lrec
        for c in range(n):
            rec = body[c * lrec:(c + 1) * lrec].decode()
            want = "%12d %21.13e" % (gid[c], vm[c])
            want += "".join(" %21.13e" % st[slot[f], c] for f in fields) + "\n"
            assert rec == want
    else:
        assert lrec == 8 * (len(fields) + 2)
        assert len(body) == n * lrec
        for c in range(n):
            w = struct.unpack("<q" + "d" * (len(fields) + 1), body[c * lrec:(c + 1) * lrec])
            assert w[0] == gid[c] and w[1] == vm[c]
            assert list(w[2:]) == [st[slot[f], c] for f in fields]
    restart = (snap / "restart").read_text()
    assert restart == ("simulate SIMULATE {\n   loop=1234; time=12.340000; "
                       "stateFile=snapshot.000000001234/state;\n}\n")
    assert os.readlink(tmp_path / "restart") == "snapshot.000000001234/restart"


@pytest.mark.parametrize("ascii", [True, False])
def test_round_trip(tmp_path, ascii):
    _lib_or_skip()
    n = 300
    gid, vm, st = _random_case(n, seed=5)
    path = hip.write_checkpoint(str(tmp_path), 7, 0.07, (40, 40, 40), gid, vm, st, ascii=ascii)
    perm = np.random.default_rng(6).permutation(n)
    vm2, st2, loop, time = hip.read_checkpoint(path, gid[perm])
    assert (loop, time) == (7, 0.07)
    if ascii:
        # 14 significant digits
        np.testing.assert_allclose(vm2, vm[perm], rtol=5.1e-14, atol=0)
        np.testing.assert_allclose(st2, st[:, perm], rtol=5.1e-14, atol=0)
    else:
        np.testing.assert_array_equal(vm2, vm[perm])
        np.testing.assert_array_equal(st2, st[:, perm])


def _write_custom(path, names, types, units, rows, binary=False, swap=False, nfiles=1,
                  per_file=None):
    """A pio state file in the reference's layout with any field list."""
    key = struct.unpack("<i", b"1234")[0]
    if swap:
        key = struct.unpack(">i", b"1234")[0]
    if binary:
        fmt = (">" if swap else "<") + "".join(
            {"u8": "q", "f8": "d", "f4": "f"}[t] for t in types)
        recs = [struct.pack(fmt, *r) for r in rows]
    else:
        recs = []
        for r in rows:
            s = " ".join(("%12d" % v) if t == "u" else ("%21.13e" % v) for v, t in zip(r, types))
            recs.append((s + "\n").encode())
    lrec = len(recs[0])
    head = ("state FILEHEADER {\n   datatype = %s;\n   nfiles = %d;\n   nrecord = %d;\n"
            "   lrec = %d;\n   endian_key = %d;\n   time = 3.500000;\n   loop = 350;\n"
            "   nfields = %d;\n   field_names = %s;\n   field_types = %s;\n"
            "   field_units = %s;\n   nx = 9; ny = 9; nz = 9;\n}\n\n") % (
        "FIXRECORDBINARY" if binary else "FIXRECORDASCII", nfiles, len(rows), lrec, key,
        len(names), " ".join(names), " ".join(types), " ".join(units))
    per_file = per_file or [len(rows)]
    k = 0
    for fid, cnt in enumerate(per_file):
        with open("%s#%06d" % (path, fid), "wb") as fh:
            if fid == 0:
                fh.write(head.encode())
            fh.write(b"".join(recs[k:k + cnt]))
        k += cnt


def test_reader_fields_by_name_units_and_files(tmp_path):
    """Field order, gid column position, unit conversion (uM -> mM, V -> mV),
    a missing field, an unknown field and a state split over two files."""
    _lib_or_skip()
    n = 50
    rng = np.random.default_rng(8)
    gid = rng.permutation(729)[:n].astype(np.int64)
    ca = rng.uniform(1e-5, 1e-3, n)        # mM
    nai = rng.uniform(5, 15, n)            # mM
    dvk = rng.uniform(-100, 100, n)        # mV
    vm = rng.uniform(-90, 30, n)
    names = ["Ca_i", "gid", "junk", "dVK_i", "Vm", "Na_i"]
    types = ["f", "u", "f", "f", "f", "f"]
    units = ["uM", "1", "1", "V", "mV", "mM"]
    rows = [(ca[c] * 1e3, gid[c], 99.0, dvk[c] * 1e-3, vm[c], nai[c]) for c in range(n)]
    base = str(tmp_path / "state")
    _write_custom(base, names, types, units, rows, nfiles=2, per_file=[20, 30])
    slot = {nm: i for i, (nm, _) in enumerate(_slot_names())}
    perm = rng.permutation(n)
    cells = gid[perm]
    import ctypes as C
    vm2 = np.zeros(n)
    st2 = np.full((hip.NSTATE, n), -7.0)
    loop, time = C.c_int(), C.c_double()
    hip.check(hip.lib().cdh_checkpoint_read(base.encode(), n, hip.ptr(cells, hip.i64p),
                                            hip.ptr(vm2), hip.ptr(st2), C.byref(loop),
                                            C.byref(time)), "read")
    np.testing.assert_allclose(st2[slot["Ca_i"]], ca[perm], rtol=1e-12)
    np.testing.assert_allclose(st2[slot["dVK_i"]], dvk[perm], rtol=1e-12)
    np.testing.assert_allclose(st2[slot["Na_i"]], nai[perm], rtol=1e-13)
    np.testing.assert_allclose(vm2, vm[perm], rtol=1e-13)
    others = [i for nm, i in slot.items() if nm not in ("Ca_i", "dVK_i", "Na_i")]
    assert (st2[others] == -7.0).all()
    assert (loop.value, time.value) == (350, 3.5)


@pytest.mark.parametrize("swap", [False, True])
def test_reader_binary_endian(tmp_path, swap):
    _lib_or_skip()
    n = 12
    rng = np.random.default_rng(9)
    gid = np.arange(n, dtype=np.int64) * 3 + 1
    h = rng.uniform(size=n)
    f4 = rng.uniform(size=n).astype(np.float32)
    rows = [(gid[c], -80.0 - c, h[c], float(f4[c])) for c in range(n)]
    base = str(tmp_path / "state")
    _write_custom(base, ["gid", "Vm", "h_gate", "m_gate"], ["u8", "f8", "f8", "f4"],
                  ["1", "mV", "1", "1"], rows, binary=True, swap=swap)
    vm, st, loop, time = hip.read_checkpoint(base, gid[::-1])
    slot = {nm: i for i, (nm, _) in enumerate(_slot_names())}
    np.testing.assert_array_equal(vm, (-80.0 - np.arange(n))[::-1])
    np.testing.assert_array_equal(st[slot["h_gate"]], h[::-1])
    np.testing.assert_array_equal(st[slot["m_gate"]], f4[::-1].astype(np.float64))


def test_reader_errors(tmp_path):
    _lib_or_skip()
    n = 10
    gid, vm, st = _random_case(n, seed=11)
    path = hip.write_checkpoint(str(tmp_path), 1, 0.01, (40, 40, 40), gid, vm, st)
    with pytest.raises(hip.CardioidHipError, match="records for"):
        hip.read_checkpoint(path, gid[:-1])
    other = gid.copy()
    other[0] = 10**9
    with pytest.raises(hip.CardioidHipError, match="which no cell has"):
        hip.read_checkpoint(path, other)
    dup = gid.copy()
    dup[1] = dup[0]
    with pytest.raises(hip.CardioidHipError, match="duplicate cell gid"):
        hip.read_checkpoint(path, dup)
    with pytest.raises(hip.CardioidHipError, match="cannot open"):
        hip.read_checkpoint(str(tmp_path / "nothing"), gid)
    base = str(tmp_path / "bad")
    _write_custom(base, ["gid", "Ca_i"], ["u", "f"], ["1", "mV"], [(int(g), 1.0) for g in gid])
    with pytest.raises(hip.CardioidHipError, match="cannot convert"):
        hip.read_checkpoint(base, gid)
