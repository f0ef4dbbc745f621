"""APD90 has no counterpart in the reference.  Its definition (oracle.h,
ora_sim_get_t90) is pinned here on synthetic action potentials whose APD90
follows in closed form."""
import numpy as np

from oracle_lib import oracle, ptr


def trace_apd(V, dt=0.5, vrest=-85.0, thr=-40.0):
    import ctypes as C
    ta, t9 = C.c_double(), C.c_double()
    V = np.ascontiguousarray(V, dtype=np.float64)
    oracle().ora_apd_trace(ptr(V), V.size, 0.0, dt, vrest, thr, C.byref(ta), C.byref(t9))
    return ta.value, t9.value


def trapezoid(t, t_up=1.0, plateau=100.0, t_end=300.0, vr=-85.0, vp=30.0):
    V = np.full(t.size, vr)
    up = (t >= t_up) & (t <= plateau)
    V[up] = vp
    dn = (t > plateau) & (t < t_end)
    V[dn] = vp - (vp - vr) * (t[dn] - plateau) / (t_end - plateau)
    return V


def test_apd90_of_a_trapezoid_action_potential():
    dt = 0.5
    t = np.arange(0, 400, dt)
    ta, t9 = trace_apd(trapezoid(t), dt)
    # V90 = -85 + 0.1*115 = -73.5, crossed at t = 100 + 200*103.5/115 = 280
    # exactly; the first sample strictly below is 280.5
    assert ta == 1.0
    assert t9 == 280.5
    assert t9 - ta == 279.5


def test_apd90_tracks_the_peak_not_the_first_sample():
    dt = 0.25
    t = np.arange(0, 400, dt)
    V = trapezoid(t, vp=10.0)
    V[(t >= 5) & (t < 6)] = 40.0          # spike above the plateau
    ta, t9 = trace_apd(V, dt)
    # with Vmax = 40 the level is -85 + 12.5 = -72.5; on the ramp from 10 to
    # -85 over 200 ms it is reached at t = 100 + 200*82.5/95
    level_t = 100 + 200 * 82.5 / 95
    assert ta == 1.0
    assert t9 == np.ceil(level_t / dt) * dt or t9 == (np.floor(level_t / dt) + 1) * dt


def test_no_activation_and_no_repolarisation():
    t = np.arange(0, 100, 0.5)
    sub = np.full(t.size, -85.0)
    sub[20:40] = -45.0                     # below the -40 mV threshold
    assert trace_apd(sub) == (-1.0, -1.0)
    stuck = np.full(t.size, -85.0)
    stuck[10:] = 20.0                      # activates, never repolarises
    ta, t9 = trace_apd(stuck)
    assert ta == 5.0 and t9 == -1.0
