"""ctypes binding of libcardioid_hip.so (the C ABI in include/cardioid_hip.h).

This module fails loudly when the HIP library is missing: there is no CPU
fallback in the product path.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

from .fitfile import DEFAULT_FIT_FILE, FitSet, read_fit_file

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcardioid_hip.so")

NSTATE = 19
MODE_FP64_UNFUSED = 0
MODE_FP64_FUSED = 1
MODE_MIXED_FUSED = 2
MODE_FP64_FMA = 3
MODES = {"fp64_unfused": MODE_FP64_UNFUSED, "fp64_fused": MODE_FP64_FUSED,
         "mixed_fused": MODE_MIXED_FUSED, "fp64_fma": MODE_FP64_FMA}

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
i64p = C.POINTER(C.c_int64)
u8p = C.POINTER(C.c_uint8)


class CellType(C.Structure):
    _fields_ = [("s_switch", C.c_int), ("P_NaK", C.c_double), ("g_Ks", C.c_double),
                ("g_Kr", C.c_double), ("g_to", C.c_double), ("g_NaL", C.c_double),
                ("Vm0", C.c_double), ("state0", C.c_double * NSTATE)]


class TT06Params(C.Structure):
    _fields_ = [("dt", C.c_double), ("tolerance", C.c_double), ("mod", C.c_int),
                ("fastGate", C.c_int), ("fastNonGate", C.c_int), ("fit_l", ip), ("fit_m", ip),
                ("fit_coef", dp), ("celltype", CellType), ("s_kr", C.c_double),
                ("s_cal", C.c_double), ("s_na", C.c_double)]


class FGRParams(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("dx", C.c_double),
                ("dy", C.c_double), ("dz", C.c_double), ("diffusionScale", C.c_double),
                ("nLocal", C.c_int64), ("nRemote", C.c_int64), ("gid", i64p), ("cellType", ip),
                ("sigma", dp)]


class BoxStim(C.Structure):
    _fields_ = [("xMin", C.c_int), ("xMax", C.c_int), ("yMin", C.c_int), ("yMax", C.c_int),
                ("zMin", C.c_int), ("zMax", C.c_int), ("t0", C.c_double), ("tf", C.c_double),
                ("period", C.c_double), ("duration", C.c_double), ("amplitude", C.c_double),
                ("phase", C.c_double)]


class SimParams(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("dx", C.c_double),
                ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
                ("diffusionScale", C.c_double), ("ncell", C.c_int64), ("gid", i64p),
                ("celltype", u8p), ("sigma", dp), ("ncelltype", C.c_int),
                ("celltypes", C.POINTER(CellType)), ("scales", dp), ("fit_l", ip), ("fit_m", ip),
                ("fit_coef", dp), ("threshold", C.c_double), ("mode", C.c_int),
                ("gate_tail_rule", C.c_int), ("sensors", C.c_int), ("variant", C.c_int),
                ("check_ranges", C.c_int), ("vm_min", C.c_double), ("vm_max", C.c_double),
                ("setup_on_host", C.c_int), ("exact_functions", C.c_int),
                ("jhTauSmooth", C.c_int)]


class BatchParams(C.Structure):
    _fields_ = [("mode", C.c_int), ("dt", C.c_double), ("fit_l", ip), ("fit_m", ip),
                ("fit_coef", dp), ("cell_capacity", C.c_int64), ("box_capacity", C.c_int64),
                ("slab_capacity", C.c_int), ("param_capacity", C.c_int), ("sensors", C.c_int),
                ("exact_functions", C.c_int), ("jhTauSmooth", C.c_int)]


class SlabDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("dx", C.c_double),
                ("dy", C.c_double), ("dz", C.c_double), ("diffusionScale", C.c_double),
                ("ncell", C.c_int64), ("gid", i64p), ("celltype", u8p), ("sigma", dp),
                ("ncelltype", C.c_int), ("celltypes", C.POINTER(CellType)), ("scales", dp),
                ("threshold", C.c_double), ("gate_tail_rule", C.c_int), ("nstim", C.c_int),
                ("stim", C.POINTER(BoxStim)), ("check_ranges", C.c_int), ("vm_min", C.c_double),
                ("vm_max", C.c_double), ("setup_on_host", C.c_int), ("nremote", C.c_int64),
                ("remote_gid", i64p), ("remote_sigma", dp)]


EXPORTS = [
    "cdh_last_error", "cdh_version", "cdh_celltype_standard",
    "cdh_tt06_create", "cdh_tt06_destroy", "cdh_tt06_calc", "cdh_tt06_update_nongate",
    "cdh_tt06_update_gate", "cdh_tt06_initialize_membrane_voltage", "cdh_tt06_checkpoint_field",
    "cdh_tt06_var_handle", "cdh_tt06_get_value", "cdh_tt06_set_value", "cdh_tt06_get_state",
    "cdh_tt06_set_state",
    "cdh_fgr_create", "cdh_fgr_destroy", "cdh_fgr_update_local_voltage",
    "cdh_fgr_update_remote_voltage", "cdh_fgr_calc", "cdh_fgr_diffusion_scale",
    "cdh_fill", "cdh_set_list", "cdh_integrate",
    "cdh_activation_sensor_create", "cdh_activation_sensor_eval", "cdh_activation_sensor_get",
    "cdh_activation_sensor_destroy", "cdh_activation_sensor_enable_apd",
    "cdh_activation_sensor_get_t90",
    "cdh_sim_create", "cdh_sim_destroy", "cdh_sim_run", "cdh_sim_sync", "cdh_sim_run_timed",
    "cdh_sim_get_vm", "cdh_sim_get_state", "cdh_sim_set_state", "cdh_sim_get_activation",
    "cdh_sim_get_t90", "cdh_sim_range_report",
    "cdh_sim_time", "cdh_sim_loop", "cdh_sim_info",
    "cdh_batch_create", "cdh_batch_destroy", "cdh_slab_extent", "cdh_batch_add",
    "cdh_batch_remove", "cdh_batch_run", "cdh_batch_run_timed", "cdh_batch_sync",
    "cdh_batch_set_graph", "cdh_batch_set_launch", "cdh_batch_get_launch", "cdh_batch_autotune",
    "cdh_batch_get_vm", "cdh_batch_get_state", "cdh_batch_get_activation", "cdh_batch_get_t90",
    "cdh_batch_range_report", "cdh_batch_slab_info", "cdh_batch_info",
    "cdh_checkpoint_write", "cdh_checkpoint_read", "cdh_sim_write_checkpoint",
    "cdh_sim_read_checkpoint", "cdh_fibre_conductivity", "cdh_fibre_conductivity_int",
    "cdh_fgr_weights", "cdh_batch_link_halo", "cdh_batch_write_checkpoint",
    "cdh_batch_read_checkpoint", "cdh_batch_get_sensors", "cdh_batch_set_sensors",
    "cdh_stream_rate", "cdh_fibre_conductivity_int_noalloc",
    "cdh_activation_sensor_counts", "cdh_sim_sensor_counts", "cdh_batch_sensor_counts",
]

ERR_FULL = 5

_lib = None


class CardioidHipError(RuntimeError):
    pass


def lib():
    """The loaded library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CardioidHipError(
                f"{LIB_PATH} is missing: build it with `python -m cardioid_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.cdh_last_error.restype = C.c_char_p
        L.cdh_fgr_diffusion_scale.restype = C.c_double
        L.cdh_sim_time.restype = C.c_double
        L.cdh_sim_time.argtypes = [C.c_void_p]
        L.cdh_sim_loop.argtypes = [C.c_void_p]
        for name in ("cdh_tt06_create",):
            getattr(L, name).argtypes = [C.POINTER(TT06Params), C.c_int, C.POINTER(C.c_void_p)]
        L.cdh_tt06_calc.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_int, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p]
        L.cdh_tt06_update_nongate.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_int,
                                              C.c_void_p, C.c_void_p, C.c_void_p]
        L.cdh_tt06_update_gate.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_int,
                                           C.c_void_p, C.c_void_p]
        L.cdh_tt06_initialize_membrane_voltage.argtypes = [C.c_void_p, C.c_void_p, C.c_int,
                                                           C.c_void_p, C.c_void_p]
        L.cdh_tt06_get_state.argtypes = [C.c_void_p, dp]
        L.cdh_tt06_set_state.argtypes = [C.c_void_p, dp]
        L.cdh_tt06_get_value.argtypes = [C.c_void_p, C.c_int, C.c_int, dp]
        L.cdh_tt06_set_value.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double]
        L.cdh_tt06_destroy.argtypes = [C.c_void_p]
        L.cdh_tt06_var_handle.argtypes = [C.c_char_p]
        L.cdh_tt06_checkpoint_field.argtypes = [C.c_int, C.POINTER(C.c_char_p),
                                                C.POINTER(C.c_char_p), ip]
        L.cdh_fgr_create.argtypes = [C.POINTER(FGRParams), C.POINTER(C.c_void_p)]
        L.cdh_fgr_destroy.argtypes = [C.c_void_p]
        L.cdh_fgr_update_local_voltage.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.cdh_fgr_update_remote_voltage.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.cdh_fgr_calc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.cdh_fgr_diffusion_scale.argtypes = [C.c_void_p]
        L.cdh_fill.argtypes = [C.c_int64, C.c_double, C.c_void_p, C.c_void_p]
        L.cdh_set_list.argtypes = [C.c_int64, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p]
        L.cdh_integrate.argtypes = [C.c_int64, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p]
        L.cdh_activation_sensor_create.argtypes = [C.c_int64, C.c_double, C.POINTER(C.c_void_p)]
        L.cdh_activation_sensor_eval.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p]
        L.cdh_activation_sensor_get.argtypes = [C.c_void_p, dp, dp, dp]
        L.cdh_activation_sensor_counts.argtypes = [C.c_void_p, i64p, i64p, i64p]
        L.cdh_sim_sensor_counts.argtypes = [C.c_void_p, i64p, i64p, i64p]
        L.cdh_batch_sensor_counts.argtypes = [C.c_void_p, C.c_int, i64p, i64p, i64p]
        L.cdh_activation_sensor_destroy.argtypes = [C.c_void_p]
        L.cdh_sim_create.argtypes = [C.POINTER(SimParams), C.c_int, C.POINTER(BoxStim),
                                     C.POINTER(C.c_void_p)]
        L.cdh_sim_destroy.argtypes = [C.c_void_p]
        L.cdh_sim_run.argtypes = [C.c_void_p, C.c_int]
        L.cdh_sim_sync.argtypes = [C.c_void_p]
        L.cdh_sim_run_timed.argtypes = [C.c_void_p, C.c_int, dp]
        L.cdh_sim_get_vm.argtypes = [C.c_void_p, dp]
        L.cdh_sim_get_state.argtypes = [C.c_void_p, dp]
        L.cdh_sim_set_state.argtypes = [C.c_void_p, dp, dp]
        L.cdh_sim_get_activation.argtypes = [C.c_void_p, dp, dp, dp]
        L.cdh_sim_get_t90.argtypes = [C.c_void_p, dp]
        L.cdh_sim_info.argtypes = [C.c_void_p, i64p, ip, ip, ip]
        L.cdh_sim_range_report.argtypes = [C.c_void_p, i64p, ip, i64p, dp, dp, dp]
        L.cdh_batch_range_report.argtypes = [C.c_void_p, C.c_int, i64p, ip, i64p, dp, dp, dp]
        L.cdh_celltype_standard.argtypes = [C.c_char_p, C.POINTER(CellType)]
        L.cdh_batch_create.argtypes = [C.POINTER(BatchParams), C.POINTER(C.c_void_p)]
        L.cdh_batch_destroy.argtypes = [C.c_void_p]
        L.cdh_slab_extent.argtypes = [C.POINTER(SlabDesc), i64p, i64p, ip]
        L.cdh_batch_add.argtypes = [C.c_void_p, C.POINTER(SlabDesc), ip]
        L.cdh_batch_remove.argtypes = [C.c_void_p, C.c_int]
        L.cdh_batch_link_halo.argtypes = [C.c_void_p, C.c_int, C.c_int, i64p]
        L.cdh_batch_write_checkpoint.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_int,
                                                 C.c_char_p, C.c_int]
        L.cdh_batch_read_checkpoint.argtypes = [C.c_void_p, C.c_int, C.c_char_p]
        L.cdh_batch_get_sensors.argtypes = [C.c_void_p, C.c_int, dp, dp, ip]
        L.cdh_batch_set_sensors.argtypes = [C.c_void_p, C.c_int, dp, C.c_double, C.c_int]
        L.cdh_batch_run.argtypes = [C.c_void_p, C.c_int]
        L.cdh_batch_run_timed.argtypes = [C.c_void_p, C.c_int, dp]
        L.cdh_batch_sync.argtypes = [C.c_void_p]
        L.cdh_batch_set_graph.argtypes = [C.c_void_p, C.c_int]
        L.cdh_batch_set_launch.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.cdh_batch_get_launch.argtypes = [C.c_void_p, ip, ip, ip, ip]
        L.cdh_batch_autotune.argtypes = [C.c_void_p, C.c_int, dp, dp, C.c_int, ip]
        L.cdh_batch_get_vm.argtypes = [C.c_void_p, C.c_int, dp]
        L.cdh_batch_get_state.argtypes = [C.c_void_p, C.c_int, dp]
        L.cdh_batch_get_activation.argtypes = [C.c_void_p, C.c_int, dp, dp, dp]
        L.cdh_batch_get_t90.argtypes = [C.c_void_p, C.c_int, dp]
        L.cdh_batch_slab_info.argtypes = [C.c_void_p, C.c_int, i64p, dp, ip]
        L.cdh_batch_info.argtypes = [C.c_void_p, i64p, i64p, i64p, ip]
        L.cdh_checkpoint_write.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_double, C.c_int,
                                           C.c_int, C.c_int, C.c_int64, i64p, dp, dp, C.c_int,
                                           C.c_char_p, C.c_int]
        L.cdh_checkpoint_read.argtypes = [C.c_char_p, C.c_int64, i64p, dp, dp, ip, dp]
        L.cdh_sim_write_checkpoint.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_char_p,
                                               C.c_int]
        L.cdh_sim_read_checkpoint.argtypes = [C.c_void_p, C.c_char_p]
        L.cdh_fibre_conductivity.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_double,
                                             C.c_double, C.c_void_p, C.c_void_p]
        L.cdh_fibre_conductivity_int.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_double,
                                                 C.c_double, C.c_void_p, C.c_void_p]
        L.cdh_fibre_conductivity_int_noalloc.argtypes = [C.c_int64, C.c_void_p, C.c_void_p,
                                                         C.c_double, C.c_double, C.c_void_p,
                                                         C.c_void_p]
        L.cdh_fgr_weights.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                      C.c_double, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
        L.cdh_stream_rate.argtypes = [C.c_int, C.c_int64, C.c_int, dp]
        _lib = L
    return _lib


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = lib().cdh_last_error().decode(errors="replace")
        raise CardioidHipError(f"{what}: error {rc}: {msg}")


def ptr(a, t=dp):
    return a.ctypes.data_as(t)


STREAM_PATTERNS = {"read": 0, "write": 1, "inplace": 2}


def stream_rate(pattern: str, cells: int, reps: int = 20) -> float:
    """GB/s of one HBM stream in the fused FP64 kernel's state pattern (19
    FP64 fields per cell in 4096-cell runs, the fields in pairs): "read",
    "write" or "inplace"
    (cdh_stream_rate)."""
    g = C.c_double()
    check(lib().cdh_stream_rate(STREAM_PATTERNS[pattern], int(cells), int(reps), C.byref(g)),
          "cdh_stream_rate")
    return g.value


def write_checkpoint(directory: str, loop: int, time: float, box: Sequence[int], gid, Vm, state,
                     ascii: bool = True, simulate_name: str = "simulate") -> str:
    """writeCheckpoint (elec/checkpointIO.cc:121-207) of host arrays: gid
    (n,), Vm (n,), state (19, n) in TT06Dev slot order.  Returns the
    stateFile name (<directory>/snapshot.<loop>/state) the restart file
    names."""
    gid = np.ascontiguousarray(gid, dtype=np.int64)
    Vm = np.ascontiguousarray(Vm, dtype=np.float64)
    state = np.ascontiguousarray(state, dtype=np.float64)
    n = gid.size
    if Vm.shape != (n,) or state.shape != (NSTATE, n):
        raise ValueError("write_checkpoint: gid (n,), Vm (n,), state (19, n) expected")
    out = C.create_string_buffer(4096)
    check(lib().cdh_checkpoint_write(directory.encode(), simulate_name.encode(), int(loop),
                                     float(time), int(box[0]), int(box[1]), int(box[2]), n,
                                     ptr(gid, i64p), ptr(Vm), ptr(state), int(bool(ascii)), out,
                                     len(out)), "cdh_checkpoint_write")
    return out.value.decode()


def read_checkpoint(state_file: str, gid):
    """readCheckpoint (elec/checkpointIO.cc:209-259): fields by name,
    records by gid.  Returns (Vm (n,), state (19, n), loop, time) in the
    order of `gid`."""
    gid = np.ascontiguousarray(gid, dtype=np.int64)
    n = gid.size
    Vm = np.empty(n)
    state = np.empty((NSTATE, n))
    loop, time = C.c_int(), C.c_double()
    check(lib().cdh_checkpoint_read(state_file.encode(), n, ptr(gid, i64p), ptr(Vm), ptr(state),
                                    C.byref(loop), C.byref(time)), "cdh_checkpoint_read")
    return Vm, state, loop.value, time.value


def _range_report(call, what):
    n, loop, cell = C.c_int64(), C.c_int(), C.c_int64()
    v, r, d = C.c_double(), C.c_double(), C.c_double()
    check(call(C.byref(n), C.byref(loop), C.byref(cell), C.byref(v), C.byref(r), C.byref(d)), what)
    if n.value == 0:
        return None
    return dict(count=n.value, loop=loop.value, cell=cell.value, V=v.value, dVmR=r.value,
                dVmD=d.value)


def standard_celltype(name: str) -> CellType:
    ct = CellType()
    check(lib().cdh_celltype_standard(name.encode(), C.byref(ct)), "cdh_celltype_standard")
    return ct


class Simulation:
    """Monodomain engine over the C ABI (cdh_sim_*).

    anatomy: an Anatomy from cardioid_amd.workloads (gid, celltype index,
    conductivity per tissue cell); stimuli: list of BoxStim.
    """

    def __init__(self, anatomy, celltypes: Sequence[str], dt: float, diffusion_scale: float,
                 stimuli=(), mode: str = "fp64_fused", threshold: float = -40.0,
                 fits: Optional[FitSet] = None, scales: Optional[np.ndarray] = None,
                 gate_tail_rule: bool = True, sensors: int = 1, variant: int = 0,
                 check_ranges: bool = True, vm_min: float = -150.0, vm_max: float = 100.0,
                 setup_on_host: bool = False, exact_functions: bool = False,
                 jhTauSmooth: int = 0):
        """exact_functions: TT06Dev configuration A (tolerance 0: the exact
        TT06 functions, generic gate update; modes fp64_unfused and
        fp64_fused, bit for bit alike), with jhTauSmooth the deck's mod
        keyword."""
        L = lib()
        self.fits = fits or read_fit_file(DEFAULT_FIT_FILE)
        fl, fm, fc = self.fits.packed()
        self._keep = [fl, fm, fc]
        cts = (CellType * len(celltypes))(*[standard_celltype(n) for n in celltypes])
        self._keep.append(cts)
        self.ncell = int(anatomy.gid.size)
        gid = np.ascontiguousarray(anatomy.gid, dtype=np.int64)
        ctype = np.ascontiguousarray(anatomy.celltype, dtype=np.uint8)
        sigma = np.ascontiguousarray(anatomy.sigma, dtype=np.float64).reshape(-1)
        self._keep += [gid, ctype, sigma]
        sc = None
        if scales is not None:
            sc = np.ascontiguousarray(scales, dtype=np.float64).reshape(-1)
            self._keep.append(sc)
        p = SimParams(nx=anatomy.nx, ny=anatomy.ny, nz=anatomy.nz, dx=anatomy.dx, dy=anatomy.dy,
                      dz=anatomy.dz, dt=dt, diffusionScale=diffusion_scale, ncell=self.ncell,
                      gid=ptr(gid, i64p), celltype=ptr(ctype, u8p), sigma=ptr(sigma),
                      ncelltype=len(celltypes), celltypes=cts,
                      scales=ptr(sc) if sc is not None else None,
                      fit_l=ptr(fl, ip), fit_m=ptr(fm, ip), fit_coef=ptr(fc), threshold=threshold,
                      mode=MODES[mode], gate_tail_rule=int(gate_tail_rule), sensors=int(sensors),
                      variant=variant, check_ranges=int(check_ranges), vm_min=vm_min,
                      vm_max=vm_max, setup_on_host=int(bool(setup_on_host)),
                      exact_functions=int(bool(exact_functions)), jhTauSmooth=int(jhTauSmooth))
        st = (BoxStim * max(1, len(stimuli)))(*stimuli)
        h = C.c_void_p()
        check(L.cdh_sim_create(C.byref(p), len(stimuli), st, C.byref(h)), "cdh_sim_create")
        self.h = h
        self.dt = dt

    def close(self):
        if getattr(self, "h", None):
            lib().cdh_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, nsteps: int):
        check(lib().cdh_sim_run(self.h, int(nsteps)), "cdh_sim_run")

    def run_timed(self, nsteps: int) -> float:
        ms = C.c_double()
        check(lib().cdh_sim_run_timed(self.h, int(nsteps), C.byref(ms)), "cdh_sim_run_timed")
        return ms.value

    def sync(self):
        check(lib().cdh_sim_sync(self.h), "cdh_sim_sync")

    def vm(self) -> np.ndarray:
        out = np.zeros(self.ncell)
        check(lib().cdh_sim_get_vm(self.h, ptr(out)), "cdh_sim_get_vm")
        return out

    def state(self) -> np.ndarray:
        out = np.zeros(NSTATE * self.ncell)
        check(lib().cdh_sim_get_state(self.h, ptr(out)), "cdh_sim_get_state")
        return out.reshape(NSTATE, self.ncell)

    def set_state(self, state=None, vm=None):
        s = None if state is None else np.ascontiguousarray(state, dtype=np.float64).reshape(-1)
        v = None if vm is None else np.ascontiguousarray(vm, dtype=np.float64)
        check(lib().cdh_sim_set_state(self.h, ptr(s) if s is not None else None,
                                      ptr(v) if v is not None else None), "cdh_sim_set_state")

    def activation(self):
        a0, a, r = (np.zeros(self.ncell) for _ in range(3))
        check(lib().cdh_sim_get_activation(self.h, ptr(a0), ptr(a), ptr(r)),
              "cdh_sim_get_activation")
        return a0, a, r

    def t90(self) -> np.ndarray:
        out = np.zeros(self.ncell)
        check(lib().cdh_sim_get_t90(self.h, ptr(out)), "cdh_sim_get_t90")
        return out

    def sensor_counts(self):
        """(cells with tact0, with tact, with trec recorded), kept on the device."""
        c = [C.c_int64() for _ in range(3)]
        check(lib().cdh_sim_sensor_counts(self.h, *(C.byref(x) for x in c)),
              "cdh_sim_sensor_counts")
        return tuple(x.value for x in c)

    def range_report(self):
        """Out-of-range Vm since the last call: None, or a dict with the
        count of (cell, step) pairs and the values Simulate::outOfRange
        prints for one of them."""
        return _range_report(lambda *a: lib().cdh_sim_range_report(self.h, *a),
                             "cdh_sim_range_report")

    def write_checkpoint(self, directory: str, ascii: bool = True) -> str:
        """snapshot.<loop>/state#000000 and restart under `directory`;
        returns the stateFile name."""
        out = C.create_string_buffer(4096)
        check(lib().cdh_sim_write_checkpoint(self.h, directory.encode(), int(bool(ascii)), out,
                                             len(out)), "cdh_sim_write_checkpoint")
        return out.value.decode()

    def read_checkpoint(self, state_file: str):
        """Restart a fresh simulation from a stateFile (states, Vm, loop
        and time)."""
        check(lib().cdh_sim_read_checkpoint(self.h, state_file.encode()),
              "cdh_sim_read_checkpoint")

    @property
    def time(self) -> float:
        return lib().cdh_sim_time(self.h)

    @property
    def loop(self) -> int:
        return lib().cdh_sim_loop(self.h)

    def info(self):
        b = C.c_int64()
        bx, by, bz = C.c_int(), C.c_int(), C.c_int()
        check(lib().cdh_sim_info(self.h, C.byref(b), C.byref(bx), C.byref(by), C.byref(bz)),
              "cdh_sim_info")
        return dict(device_bytes=b.value, box=(bx.value, by.value, bz.value))


class Slab:
    """One slab description for the batched engine (cdh_slab_desc): an
    anatomy with its cell types, stimuli and per-cell drug scales."""

    def __init__(self, anatomy, celltypes: Sequence[str], diffusion_scale: float, stimuli=(),
                 scales=None, threshold: float = -40.0, gate_tail_rule: bool = True,
                 check_ranges: bool = True, vm_min: float = -150.0, vm_max: float = 100.0,
                 setup_on_host: bool = False, remote_gid=None, remote_sigma=None):
        cts = (CellType * len(celltypes))(*[standard_celltype(n) for n in celltypes])
        self.ncell = int(anatomy.gid.size)
        gid = np.ascontiguousarray(anatomy.gid, dtype=np.int64)
        ctype = np.ascontiguousarray(anatomy.celltype, dtype=np.uint8)
        sigma = np.ascontiguousarray(anatomy.sigma, dtype=np.float64).reshape(-1)
        sc = None
        if scales is not None:
            sc = np.asarray(scales, dtype=np.float64)
            if sc.size == 3:
                sc = np.tile(sc, self.ncell)
            sc = np.ascontiguousarray(sc.reshape(-1))
        st = (BoxStim * max(1, len(stimuli)))(*stimuli)
        rg = np.ascontiguousarray(remote_gid if remote_gid is not None else [], dtype=np.int64)
        rs = np.ascontiguousarray(remote_sigma if remote_sigma is not None else np.zeros((0, 6)),
                                  dtype=np.float64).reshape(-1)
        if rs.size != 6 * rg.size:
            raise ValueError("remote_sigma must hold 6 values per remote gid")
        self.nremote = int(rg.size)
        self._keep = [cts, gid, ctype, sigma, sc, st, rg, rs]
        self.desc = SlabDesc(nx=anatomy.nx, ny=anatomy.ny, nz=anatomy.nz, dx=anatomy.dx,
                             dy=anatomy.dy, dz=anatomy.dz, diffusionScale=diffusion_scale,
                             ncell=self.ncell, gid=ptr(gid, i64p), celltype=ptr(ctype, u8p),
                             sigma=ptr(sigma), ncelltype=len(celltypes), celltypes=cts,
                             scales=ptr(sc) if sc is not None else None, threshold=threshold,
                             gate_tail_rule=int(gate_tail_rule), nstim=len(stimuli), stim=st,
                             check_ranges=int(check_ranges), vm_min=vm_min, vm_max=vm_max,
                             setup_on_host=int(bool(setup_on_host)), nremote=rg.size,
                             remote_gid=ptr(rg, i64p) if rg.size else None,
                             remote_sigma=ptr(rs) if rs.size else None)

    def extent(self):
        """(cells, box points, parameter sets) the slab takes in an arena."""
        c, b, p = C.c_int64(), C.c_int64(), C.c_int()
        check(lib().cdh_slab_extent(C.byref(self.desc), C.byref(c), C.byref(b), C.byref(p)),
              "cdh_slab_extent")
        return c.value, b.value, p.value


class ArenaFull(CardioidHipError):
    pass


class Batch:
    """Batched engine over the C ABI (cdh_batch_*): slabs advanced together,
    one launch per step, optionally replayed from captured HIP graphs."""

    def __init__(self, dt: float, cell_capacity: int, box_capacity: int, slab_capacity: int = 256,
                 param_capacity: int = 4096, mode: str = "mixed_fused", sensors: int = 1,
                 fits: Optional[FitSet] = None, exact_functions: bool = False,
                 jhTauSmooth: int = 0):
        """exact_functions: TT06Dev configuration A (tolerance 0, the reaction
        factory's default deck) in mode fp64_fused; jhTauSmooth the deck's mod."""
        self.fits = fits or read_fit_file(DEFAULT_FIT_FILE)
        fl, fm, fc = self.fits.packed()
        self._keep = [fl, fm, fc]
        p = BatchParams(mode=MODES[mode], dt=dt, fit_l=ptr(fl, ip), fit_m=ptr(fm, ip),
                        fit_coef=ptr(fc), cell_capacity=int(cell_capacity),
                        box_capacity=int(box_capacity), slab_capacity=int(slab_capacity),
                        param_capacity=int(param_capacity), sensors=int(sensors),
                        exact_functions=int(bool(exact_functions)), jhTauSmooth=int(jhTauSmooth))
        h = C.c_void_p()
        check(lib().cdh_batch_create(C.byref(p), C.byref(h)), "cdh_batch_create")
        self.h = h
        self.dt = dt
        self.ncell = {}

    def close(self):
        if getattr(self, "h", None):
            lib().cdh_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, slab: Slab) -> int:
        sid = C.c_int()
        rc = lib().cdh_batch_add(self.h, C.byref(slab.desc), C.byref(sid))
        if rc == ERR_FULL:
            raise ArenaFull("arena full")
        check(rc, "cdh_batch_add")
        self.ncell[sid.value] = slab.ncell
        return sid.value

    def remove(self, sid: int):
        check(lib().cdh_batch_remove(self.h, sid), "cdh_batch_remove")
        self.ncell.pop(sid, None)

    def write_checkpoint(self, sid: int, directory: str, ascii: bool = True) -> str:
        """snapshot.<loop>/state#000000 and restart of one slab; returns the
        stateFile name."""
        out = C.create_string_buffer(4096)
        check(lib().cdh_batch_write_checkpoint(self.h, sid, directory.encode(), int(bool(ascii)),
                                               out, len(out)), "cdh_batch_write_checkpoint")
        return out.value.decode()

    def read_checkpoint(self, sid: int, state_file: str):
        """States, Vm and clock of slab `sid` from a stateFile."""
        check(lib().cdh_batch_read_checkpoint(self.h, sid, state_file.encode()),
              "cdh_batch_read_checkpoint")

    def sensors(self, sid: int):
        """(rows (6, ncell): tact0, tact, trec, Vmax, t90, APD phase; time,
        loop) of slab sid, the run state a checkpoint file does not hold."""
        rows = np.zeros(6 * self.ncell[sid])
        t, l = C.c_double(), C.c_int()
        check(lib().cdh_batch_get_sensors(self.h, sid, ptr(rows), C.byref(t), C.byref(l)),
              "cdh_batch_get_sensors")
        return rows.reshape(6, -1), t.value, l.value

    def set_sensors(self, sid: int, rows, time: float, loop: int):
        """Restore what sensors() returned: the sensors then continue the run."""
        r = np.ascontiguousarray(rows, dtype=np.float64).reshape(-1)
        if r.size != 6 * self.ncell[sid]:
            raise ValueError("set_sensors: 6 rows of ncell values expected")
        check(lib().cdh_batch_set_sensors(self.h, sid, ptr(r), float(time), int(loop)),
              "cdh_batch_set_sensors")

    def link_halo(self, dst: int, src: int) -> int:
        """After every step, slab dst's remote cells owned by slab src take
        src's Vm (HaloExchange on one device); returns the points linked."""
        n = C.c_int64()
        check(lib().cdh_batch_link_halo(self.h, dst, src, C.byref(n)), "cdh_batch_link_halo")
        return n.value

    def run(self, nsteps: int):
        check(lib().cdh_batch_run(self.h, int(nsteps)), "cdh_batch_run")

    def run_timed(self, nsteps: int) -> float:
        ms = C.c_double()
        check(lib().cdh_batch_run_timed(self.h, int(nsteps), C.byref(ms)), "cdh_batch_run_timed")
        return ms.value

    def sync(self):
        check(lib().cdh_batch_sync(self.h), "cdh_batch_sync")

    def set_graph(self, steps: int):
        check(lib().cdh_batch_set_graph(self.h, int(steps)), "cdh_batch_set_graph")

    def set_launch(self, waves: int, block: int, occupancy: int = 0, tile_order: int = -1):
        check(lib().cdh_batch_set_launch(self.h, waves, block, occupancy, tile_order),
              "cdh_batch_set_launch")

    def launch(self):
        """(register target, workgroup size, occupancy cap, tile order)"""
        w, b, o, x = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(lib().cdh_batch_get_launch(self.h, C.byref(w), C.byref(b), C.byref(o), C.byref(x)),
              "cdh_batch_get_launch")
        return w.value, b.value, o.value, x.value

    def autotune(self, steps: int = 10):
        """(best ms per step, [(waves, block, occupancy, tile order, ms per
        step), ...])"""
        best = C.c_double()
        table = np.zeros(5 * 64)
        nrows = C.c_int()
        check(lib().cdh_batch_autotune(self.h, int(steps), C.byref(best), ptr(table), 64,
                                       C.byref(nrows)), "cdh_batch_autotune")
        rows = [(int(table[5 * k]), int(table[5 * k + 1]), int(table[5 * k + 2]),
                 int(table[5 * k + 3]), float(table[5 * k + 4]))
                for k in range(min(nrows.value, 64))]
        return best.value, rows

    def vm(self, sid: int) -> np.ndarray:
        out = np.zeros(self.ncell[sid])
        check(lib().cdh_batch_get_vm(self.h, sid, ptr(out)), "cdh_batch_get_vm")
        return out

    def state(self, sid: int) -> np.ndarray:
        out = np.zeros(NSTATE * self.ncell[sid])
        check(lib().cdh_batch_get_state(self.h, sid, ptr(out)), "cdh_batch_get_state")
        return out.reshape(NSTATE, -1)

    def activation(self, sid: int):
        a0, a, r = (np.zeros(self.ncell[sid]) for _ in range(3))
        check(lib().cdh_batch_get_activation(self.h, sid, ptr(a0), ptr(a), ptr(r)),
              "cdh_batch_get_activation")
        return a0, a, r

    def t90(self, sid: int) -> np.ndarray:
        out = np.zeros(self.ncell[sid])
        check(lib().cdh_batch_get_t90(self.h, sid, ptr(out)), "cdh_batch_get_t90")
        return out

    def sensor_counts(self, sid: int):
        """(cells with tact0, with tact, with trec recorded) in slab sid."""
        c = [C.c_int64() for _ in range(3)]
        check(lib().cdh_batch_sensor_counts(self.h, sid, *(C.byref(x) for x in c)),
              "cdh_batch_sensor_counts")
        return tuple(x.value for x in c)

    def range_report(self, sid: int):
        return _range_report(lambda *a: lib().cdh_batch_range_report(self.h, sid, *a),
                             "cdh_batch_range_report")

    def slab_info(self, sid: int):
        n, t, l = C.c_int64(), C.c_double(), C.c_int()
        check(lib().cdh_batch_slab_info(self.h, sid, C.byref(n), C.byref(t), C.byref(l)),
              "cdh_batch_slab_info")
        return dict(ncell=n.value, time=t.value, loop=l.value)

    def info(self):
        b, c, x, n = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int()
        check(lib().cdh_batch_info(self.h, C.byref(b), C.byref(c), C.byref(x), C.byref(n)),
              "cdh_batch_info")
        return dict(device_bytes=b.value, cells_used=c.value, box_used=x.value, nslab=n.value)
