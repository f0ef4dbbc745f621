"""ctypes bindings to the CPU oracle (oracle/liboracle.so) and, where it was
built, to oracle/_ref/libcardioid_ref.so (the reference's own kernels).

Test infrastructure only.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libcardioid_ref.so")

NSTATE = 19
STATE_NAMES = ["Ca_i", "dVK_i", "Na_i", "Ca_ss", "Ca_SR", "R_prime", "fCass", "m", "h", "j",
               "Xr1", "Xr2", "Xs", "r", "d", "f", "f2", "jL", "s"]

dp = C.POINTER(C.c_double)
fp = C.POINTER(C.c_float)
i64p = C.POINTER(C.c_int64)
u8p = C.POINTER(C.c_uint8)
ip = C.POINTER(C.c_int)


class OraPade(C.Structure):
    _fields_ = [("l", C.c_int), ("m", C.c_int), ("coef", C.c_double * 40)]


class OraFitset(C.Structure):
    _fields_ = [("fit", OraPade * 33), ("dt_scaled", C.c_int)]


class OraCelltype(C.Structure):
    _fields_ = [("s_switch", C.c_int), ("P_NaK", C.c_double), ("g_Ks", C.c_double),
                ("g_Kr", C.c_double), ("g_to", C.c_double), ("g_NaL", C.c_double),
                ("Vm0", C.c_double), ("state0", C.c_double * NSTATE)]


class OraScales(C.Structure):
    _fields_ = [("kr", C.c_double), ("cal", C.c_double), ("na", C.c_double)]


class OraSimParams(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("dt", C.c_double), ("diffusionScale", C.c_double),
                ("ncell", C.c_int64), ("gid", i64p), ("celltype", u8p), ("sigma", dp),
                ("ncelltype", C.c_int), ("celltypes", C.POINTER(OraCelltype)),
                ("scales", C.POINTER(OraScales)), ("fits", C.POINTER(OraFitset)),
                ("exact", C.c_int), ("threshold", C.c_double), ("nthreads", C.c_int),
                ("jhTauSmooth", C.c_int), ("no_loop_bounds", C.c_int)]


class OraBoxStim(C.Structure):
    _fields_ = [("xMin", C.c_int), ("xMax", C.c_int), ("yMin", C.c_int), ("yMax", C.c_int),
                ("zMin", C.c_int), ("zMax", C.c_int), ("t0", C.c_double), ("tf", C.c_double),
                ("period", C.c_double), ("duration", C.c_double), ("amplitude", C.c_double),
                ("phase", C.c_double)]


def ptr(a, t=dp):
    return a.ctypes.data_as(t)


_ora = None
_ref = None


def oracle():
    global _ora
    if _ora is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle`")
        L = C.CDLL(ORACLE_SO)
        L.ora_c9.restype = C.c_double
        L.ora_pade_eval.restype = C.c_double
        L.ora_pade_eval.argtypes = [C.POINTER(OraPade), C.c_double]
        L.ora_tt06_exact_func.restype = C.c_double
        L.ora_tt06_exact_func.argtypes = [C.c_int, C.c_double, C.c_int]
        L.ora_tt06_mod_coefficients.argtypes = [C.c_int, dp]
        L.ora_celltype_standard.argtypes = [C.c_char_p, C.POINTER(OraCelltype)]
        L.ora_fitset_scale_dt.argtypes = [C.POINTER(OraFitset), C.c_double]
        L.ora_tt06_nongate.argtypes = [C.POINTER(OraFitset), C.c_double, C.POINTER(OraCelltype),
                                       C.POINTER(OraScales), C.c_int, dp, dp, C.c_int64, dp]
        L.ora_tt06_gates_v2.argtypes = [C.POINTER(OraFitset), C.c_int, C.c_int, dp, dp, C.c_int64]
        L.ora_tt06_gate_v2.argtypes = [C.POINTER(OraFitset), C.c_int, C.c_int, dp, dp]
        L.ora_tt06_nongate_exact.argtypes = [C.c_double, C.POINTER(OraCelltype),
                                             C.POINTER(OraScales), C.c_int, dp, dp, C.c_int64, dp]
        L.ora_tt06_gates_exact.argtypes = [C.c_double, C.c_int, C.c_int, C.c_int, dp, dp, C.c_int64]
        L.ora_gate_flags.argtypes = [C.c_int64, C.c_int64]
        L.ora_tt06_init_state.argtypes = [C.POINTER(OraCelltype), C.c_int, dp, C.c_int64, dp]
        L.ora_fgr_offsets.argtypes = [C.c_int, C.c_int, C.c_int, i64p]
        L.ora_fgr_weights.argtypes = [C.c_int, C.c_int, C.c_int, u8p, dp, C.c_double, C.c_double,
                                      C.c_double, C.c_int64, i64p, fp, dp]
        L.ora_fgr_f2.argtypes = [C.c_uint, ip, dp]
        L.ora_fibre_conductivity.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, dp]
        L.ora_fgr_calc.argtypes = [C.c_int64, i64p, fp, dp, i64p, dp, C.c_double, dp]
        L.ora_sim_create.restype = C.c_void_p
        L.ora_sim_create.argtypes = [C.POINTER(OraSimParams), C.c_int, C.POINTER(OraBoxStim)]
        L.ora_sim_destroy.argtypes = [C.c_void_p]
        L.ora_sim_run.argtypes = [C.c_void_p, C.c_int]
        for fn in ("ora_sim_get_vm", "ora_sim_get_state", "ora_sim_get_activation_vm0",
                   "ora_sim_get_activation_thr", "ora_sim_get_recovery_thr"):
            getattr(L, fn).argtypes = [C.c_void_p, dp]
        L.ora_sim_set_state.argtypes = [C.c_void_p, dp, dp]
        L.ora_sim_restart.argtypes = [C.c_void_p, dp, dp, C.c_int, C.c_double]
        L.ora_sim_set_reaction.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int]
        L.ora_sim_get_t90.argtypes = [C.c_void_p, dp]
        L.ora_apd_trace.argtypes = [dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                    dp, dp]
        L.ora_sim_time.restype = C.c_double
        L.ora_sim_time.argtypes = [C.c_void_p]
        L.ora_sim_loop.argtypes = [C.c_void_p]
        L.ora_tt06_init()
        _ora = L
    return _ora


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        L = C.CDLL(REF_SO)
        L.ref_get_c9.restype = C.c_double
        L.ref_load_fits.argtypes = [ip, ip, dp, C.c_double, C.c_int, C.c_double]
        L.ref_fit_eval.restype = C.c_double
        L.ref_fit_eval.argtypes = [C.c_int, C.c_double]
        L.ref_set_celltype.argtypes = [C.c_int] + [C.c_double] * 5
        L.ref_nongate.argtypes = [C.c_double, C.c_int, C.c_long, dp, dp, dp]
        L.ref_gates_fast.argtypes = [C.c_double, C.c_int, C.c_long, dp, dp]
        L.ref_gates_generic.argtypes = [C.c_double, C.c_int, C.c_long, dp, dp]
        L.ref_gate_v2.argtypes = [C.c_double, C.c_int, C.c_int, dp, dp]
        L.ref_exact_func.restype = C.c_double
        L.ref_exact_func.argtypes = [C.c_int, C.c_double, C.c_int]
        L.ref_fgr_f2.argtypes = [C.c_uint, ip, dp]
        L.ref_celltype_standard.argtypes = [C.c_char_p, ip, dp, dp]
        L.ref_set_group_celltype.argtypes = [C.c_int, C.c_int] + [C.c_double] * 5
        L.ref_celltype_names.argtypes = [C.c_char_p, C.c_int]
        L.ref_fibre_conductivity.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, dp]
        L.ref_init()
        _ref = L
    return _ref


def ref_celltype_table():
    """TT06Func::getStandardCellTypes of the reference build: name ->
    (s_switch, [P_NaK, g_Ks, g_Kr, g_to, g_NaL, Vm], state[19] in slot order
    with K_i in slot 1)."""
    R = ref()
    buf = C.create_string_buffer(4096)
    n = R.ref_celltype_names(buf, len(buf))
    names = buf.value.decode().split()
    assert n == len(names)
    out = {}
    for name in names:
        sw = C.c_int()
        parms, state = np.zeros(6), np.full(NSTATE, np.nan)
        cnt = R.ref_celltype_standard(name.encode(), C.byref(sw), ptr(parms), ptr(state))
        assert cnt == NSTATE, (name, cnt)
        out[name] = (sw.value, parms, state)
    return out


def aligned(n, dtype=np.float64, align=64):
    """Zeroed array whose data pointer is `align`-byte aligned (the v2 gate
    kernels load 4 doubles from 32-byte aligned addresses)."""
    dtype = np.dtype(dtype)
    buf = np.zeros(n * dtype.itemsize + align, dtype=np.uint8)
    off = (-buf.ctypes.data) % align
    return buf[off:off + n * dtype.itemsize].view(dtype)


def make_fitset(fs, dt=None) -> OraFitset:
    """ctypes fit set from a cardioid_amd.fitfile.FitSet, optionally dt-scaled."""
    out = OraFitset()
    for i, f in enumerate(fs.fits):
        out.fit[i].l = f.l
        out.fit[i].m = f.m
        for k, c in enumerate(f.coef):
            out.fit[i].coef[k] = c
    out.dt_scaled = 0
    if dt is not None:
        oracle().ora_fitset_scale_dt(C.byref(out), dt)
    return out


def celltype(name: str) -> OraCelltype:
    ct = OraCelltype()
    rc = oracle().ora_celltype_standard(name.encode(), C.byref(ct))
    if rc != 0:
        raise KeyError(name)
    return ct


def ref_load(fs, dt, exact=False, jhTauSmooth=1):
    l, m, coef = fs.packed()
    R = ref()
    R.ref_load_fits(ptr(l, ip), ptr(m, ip), ptr(coef), 0.0 if exact else 1e-4,
                    jhTauSmooth, dt if dt else 0.0)
    return R


class OracleSim:
    """The oracle's omp-loop simulation on an Anatomy (cardioid_amd.workloads)."""

    def __init__(self, anat, celltypes, dt, diffusion_scale, stimuli=(), exact=False,
                 threshold=-40.0, fits=None, scales=None, nthreads=0, jhTauSmooth=0,
                 gate_tail_rule=True):
        from cardioid_amd.fitfile import read_fit_file
        O = oracle()
        fs = fits or read_fit_file()
        self.fits, self.dt = fs, dt
        self.ofs = make_fitset(fs)
        self.cts = (OraCelltype * len(celltypes))(*[celltype(n) for n in celltypes])
        self.gid = np.ascontiguousarray(anat.gid, dtype=np.int64)
        self.ct = np.ascontiguousarray(anat.celltype, dtype=np.uint8)
        self.sig = np.ascontiguousarray(anat.sigma, dtype=np.float64).reshape(-1)
        self.sc = None
        if scales is not None:
            s = np.ascontiguousarray(scales, dtype=np.float64).reshape(-1, 3)
            self.sc = (OraScales * s.shape[0])(*[OraScales(*row) for row in s])
        if nthreads <= 0:
            nthreads = min(16, os.cpu_count() or 1)
        p = OraSimParams(nx=anat.nx, ny=anat.ny, nz=anat.nz, dx=anat.dx, dy=anat.dy, dz=anat.dz,
                         dt=dt, diffusionScale=diffusion_scale, ncell=self.gid.size,
                         gid=ptr(self.gid, i64p), celltype=ptr(self.ct, u8p), sigma=ptr(self.sig),
                         ncelltype=len(celltypes), celltypes=self.cts, scales=self.sc,
                         fits=C.pointer(self.ofs), exact=int(exact), threshold=threshold,
                         nthreads=nthreads, jhTauSmooth=int(jhTauSmooth),
                         no_loop_bounds=int(not gate_tail_rule))
        st = (OraBoxStim * max(1, len(stimuli)))()
        for k, b in enumerate(stimuli):
            for f, _ in OraBoxStim._fields_:
                setattr(st[k], f, getattr(b, f))
        self.h = O.ora_sim_create(C.byref(p), len(stimuli), st)
        self.n = self.gid.size

    def run(self, nsteps):
        oracle().ora_sim_run(self.h, nsteps)

    def vm(self):
        out = np.zeros(self.n)
        oracle().ora_sim_get_vm(self.h, ptr(out))
        return out

    def state(self):
        out = np.zeros(NSTATE * self.n)
        oracle().ora_sim_get_state(self.h, ptr(out))
        return out.reshape(NSTATE, self.n)

    def use_reference_reaction(self, react_threads=1, pointwise_threads=1):
        """Run the reaction through the reference's own kernels
        (update_nonGate and the v2 gate kernels, compiled from source into
        oracle/_ref) instead of the oracle's restatement: react_threads = 1
        is the reference's one thread per reaction group, pointwise_threads =
        1 its serial integrator and sensor loops.  Configuration B only."""
        R = ref_load(self.fits, self.dt, exact=False, jhTauSmooth=1)
        for g, ct in enumerate(self.cts):
            R.ref_set_group_celltype(g, ct.s_switch, ct.P_NaK, ct.g_Ks, ct.g_Kr, ct.g_to,
                                     ct.g_NaL)
        fn = C.cast(R.ref_react_group, C.c_void_p)
        oracle().ora_sim_set_reaction(self.h, fn, None, int(react_threads),
                                      int(pointwise_threads))

    def restart(self, state, vm, loop, time):
        """States [19][n] and Vm, the clock, and fresh sensors, as the
        reference restarts from a checkpoint."""
        st = np.ascontiguousarray(state, dtype=np.float64).reshape(-1)
        v = np.ascontiguousarray(vm, dtype=np.float64)
        oracle().ora_sim_restart(self.h, ptr(st), ptr(v), int(loop), float(time))

    def activation(self):
        a0, a, r = (np.zeros(self.n) for _ in range(3))
        O = oracle()
        O.ora_sim_get_activation_vm0(self.h, ptr(a0))
        O.ora_sim_get_activation_thr(self.h, ptr(a))
        O.ora_sim_get_recovery_thr(self.h, ptr(r))
        return a0, a, r

    def t90(self):
        out = np.zeros(self.n)
        oracle().ora_sim_get_t90(self.h, ptr(out))
        return out

    @property
    def time(self):
        return oracle().ora_sim_time(self.h)

    def close(self):
        if self.h:
            oracle().ora_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
