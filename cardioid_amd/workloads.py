"""Synthetic anatomies, conductivities and stimuli for the benchmark and
parity workloads (seeded; no input files).

Units follow Cardioid's internal ones (elec/cardioid.cc:43-63): mm, ms, mV,
conductivity in S/mm, diffusionScale in mm^3/mF.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Tuple

import numpy as np

# Niederer-deck diffusion scale (test/tests/Niederer/object.data, fgr DIFFUSION)
DIFFUSION_SCALE = 714.2857143
SIGMA_L = 0.1334e-3    # S/mm  (0.1334 mS/mm)
SIGMA_T = 0.0176e-3    # S/mm  (0.0176 mS/mm)
CELLTYPES_EMI = ["endoCellML", "midCellML", "epiCellML"]


@dataclass
class Anatomy:
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    gid: np.ndarray        # int64, ascending; gid = x + nx*(y + ny*z)
    celltype: np.ndarray   # uint8 index into the cell-type list
    sigma: np.ndarray      # (n, 6): a11 a12 a13 a22 a23 a33, S/mm

    @property
    def ncell(self) -> int:
        return int(self.gid.size)

    def tuples(self):
        g = self.gid
        x = g % self.nx
        y = (g // self.nx) % self.ny
        z = g // (self.nx * self.ny)
        return x, y, z


def fibre_tensor(f: np.ndarray, sl: float = SIGMA_L, st: float = SIGMA_T) -> np.ndarray:
    """sigma = st*I + (sl - st) f f^T for unit fibre vectors f (n, 3), the
    tensor FibreConductivity builds from the fibre angles
    (elec/FibreConductivity.cc:27-62)."""
    d = sl - st
    out = np.empty((f.shape[0], 6))
    out[:, 0] = st + d * f[:, 0] * f[:, 0]
    out[:, 1] = d * f[:, 0] * f[:, 1]
    out[:, 2] = d * f[:, 0] * f[:, 2]
    out[:, 3] = st + d * f[:, 1] * f[:, 1]
    out[:, 4] = d * f[:, 1] * f[:, 2]
    out[:, 5] = st + d * f[:, 2] * f[:, 2]
    return out


def brick(nx: int, ny: int, nz: int, dx: float = 0.2, celltype: int = 0,
          sigma=(SIGMA_L, SIGMA_T, SIGMA_T)) -> Anatomy:
    """All-tissue brick with uniform diagonal conductivity (the `uniform`
    CONDUCTIVITY method, elec/setConductivity.cc:234-247)."""
    n = nx * ny * nz
    gid = np.arange(n, dtype=np.int64)
    s = np.zeros((n, 6))
    s[:, 0], s[:, 3], s[:, 5] = sigma
    return Anatomy(nx, ny, nz, dx, dx, dx, gid, np.full(n, celltype, np.uint8), s)


def ellipsoid_shell(nx: int = 256, ny: int = 256, nz: int = 512, dx: float = 0.2, seed: int = 0,
                    occupancy: float = 0.60, wall: float = 0.45) -> Tuple[Anatomy, np.ndarray]:
    """Truncated ellipsoidal shell in the (nx, ny, nz) box.

    The outer surface is the ellipsoid rho = 1 with semi-axes s*(nx, ny, nz)/2
    (clipped by the box), the inner one rho = 1 - wall; s is chosen so the
    tissue fills `occupancy` of the box.  Cell types by normalised transmural
    depth d (0 at the outer, epicardial surface, 1 at the inner one): epi for
    d < 1/3, mid for 1/3 <= d < 2/3, endo beyond.  Fibres lie in the surface
    tangent plane at helix angle -60 deg (epi) to +60 deg (endo), linear in d,
    with a seeded +-2 deg jitter.  Returns the anatomy and the depth d.
    """
    rng = np.random.default_rng(seed)
    cx, cy, cz = (nx - 1) / 2, (ny - 1) / 2, (nz - 1) / 2

    def occ(s):
        xs = (np.arange(nx) - cx) / (s * nx / 2)
        ys = (np.arange(ny) - cy) / (s * ny / 2)
        zs = (np.arange(nz) - cz) / (s * nz / 2)
        r2 = xs[:, None, None] ** 2 + ys[None, :, None] ** 2  # (x, y, 1)
        frac = 0.0
        for k in range(0, nz, max(1, nz // 64)):
            rr = np.sqrt(r2[:, :, 0] + zs[k] ** 2)
            frac += np.mean((rr <= 1.0) & (rr >= 1.0 - wall))
        return frac / len(range(0, nz, max(1, nz // 64)))

    lo, hi = 0.5, 2.0
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        if occ(mid) < occupancy:
            lo = mid
        else:
            hi = mid
    s = 0.5 * (lo + hi)
    ax, ay, az = s * nx / 2, s * ny / 2, s * nz / 2

    gids, depth, fib = [], [], []
    xs = np.arange(nx)
    X, Y = np.meshgrid(xs, np.arange(ny), indexing="ij")   # (x, y)
    ex, ey = (X - cx) / ax, (Y - cy) / ay
    for z in range(nz):
        ez = (z - cz) / az
        rho = np.sqrt(ex ** 2 + ey ** 2 + ez ** 2)
        m = (rho <= 1.0) & (rho >= 1.0 - wall)
        if not m.any():
            continue
        xx, yy = X[m], Y[m]
        d = (1.0 - rho[m]) / wall
        # surface frame: outward normal of the ellipsoid, circumferential
        # direction around z, longitudinal direction in the tangent plane
        nrm = np.stack([ex[m] / ax, ey[m] / ay, np.full(xx.size, ez / az)], -1)
        nrm /= np.linalg.norm(nrm, axis=1, keepdims=True) + 1e-30
        circ = np.stack([-(yy - cy), (xx - cx), np.zeros(xx.size)], -1).astype(np.float64)
        cn = np.linalg.norm(circ, axis=1, keepdims=True)
        circ = np.where(cn > 1e-9, circ / np.maximum(cn, 1e-30), np.array([1.0, 0.0, 0.0]))
        lon = np.cross(nrm, circ)
        lon /= np.linalg.norm(lon, axis=1, keepdims=True) + 1e-30
        ang = np.deg2rad(-60.0 + 120.0 * d + rng.uniform(-2, 2, d.size))
        f = np.cos(ang)[:, None] * circ + np.sin(ang)[:, None] * lon
        f /= np.linalg.norm(f, axis=1, keepdims=True)
        gids.append(xx + nx * (yy + ny * z))
        depth.append(d)
        fib.append(f)
    gid = np.concatenate(gids).astype(np.int64)
    order = np.argsort(gid, kind="stable")
    gid = gid[order]
    depth = np.concatenate(depth)[order]
    f = np.concatenate(fib)[order]
    ctype = np.where(depth < 1 / 3, 2, np.where(depth < 2 / 3, 1, 0)).astype(np.uint8)
    return Anatomy(nx, ny, nz, dx, dx, dx, gid, ctype, fibre_tensor(f)), depth


def box_stimulus(xMin=-1, xMax=None, yMin=-1, yMax=None, zMin=-1, zMax=None, nx=0, ny=0, nz=0,
                 vStim=-52.0, tStart=0.0, duration=1.0, period=1000.0, t0=-1000.0, tf=1e30):
    """A box STIMULUS with a periodic pulse and the factory's defaults
    (elec/stimulusFactory.cc:24-67, :93-118): amplitude = -vStim."""
    from .hip import BoxStim
    return BoxStim(xMin=xMin, xMax=nx if xMax is None else xMax, yMin=yMin,
                   yMax=ny if yMax is None else yMax, zMin=zMin,
                   zMax=nz if zMax is None else zMax, t0=t0, tf=tf, period=period,
                   duration=duration, amplitude=-vStim, phase=tStart)


def endocardial_point_stimuli(anat: Anatomy, depth: np.ndarray, n: int = 8, seed: int = 0,
                              half: int = 2, vStim: float = -52.0, duration: float = 1.0):
    """n seeded sites on the inner (endocardial) surface; each a small box
    stimulus of (2*half-1)^3 grid points around the site."""
    rng = np.random.default_rng(seed)
    cand = np.nonzero(depth > 0.97)[0]
    pick = rng.choice(cand, size=n, replace=False)
    x, y, z = anat.tuples()
    out: List = []
    for c in pick:
        out.append(box_stimulus(x[c] - half, x[c] + half, y[c] - half, y[c] + half,
                                z[c] - half, z[c] + half, vStim=vStim, duration=duration))
    return out


def fibre_slab(nx: int, ny: int, nz: int, phi_deg: float, dx: float = 0.2,
               celltype: int = 0) -> Anatomy:
    """All-tissue slab with one fibre direction in the x-y plane at angle phi
    from x (the `fibre` conductivity with theta = 90 deg,
    elec/FibreConductivity.cc:19-62)."""
    n = nx * ny * nz
    p = np.deg2rad(phi_deg)
    f = np.tile(np.array([np.cos(p), np.sin(p), 0.0]), (n, 1))
    return Anatomy(nx, ny, nz, dx, dx, dx, np.arange(n, dtype=np.int64),
                   np.full(n, celltype, np.uint8), fibre_tensor(f))


@dataclass
class SlabCase:
    anatomy: Anatomy
    scales: np.ndarray      # (3,): s_kr, s_cal, s_na
    stimuli: list
    phi_deg: float


def _log_uniform_dims(rng, lo, hi):
    return tuple(int(round(np.exp(rng.uniform(np.log(a), np.log(b))))) for a, b in zip(lo, hi))


def batched_slabs(n: int = 64, seed: int = 1, lo=(32, 32, 32), hi=(160, 160, 96)) -> List[SlabCase]:
    """Independent tissue slabs for the batched engine: each axis drawn
    log-uniform between lo and hi, drug scales (s_kr, s_cal, s_na) uniform in
    [0, 1], fibre angle uniform in [-90, 90] deg, endocardial cells, and the
    box stimulus x < 4 (-52 uA/uF for 1 ms at t = 0)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nx, ny, nz = _log_uniform_dims(rng, lo, hi)
        phi = float(rng.uniform(-90.0, 90.0))
        sc = rng.uniform(0.0, 1.0, 3)
        a = fibre_slab(nx, ny, nz, phi)
        out.append(SlabCase(a, sc, [box_stimulus(xMax=4, nx=nx, ny=ny, nz=nz)], phi))
    return out


@dataclass
class Request:
    arrival_s: float        # wall-clock arrival after the stream starts
    dims: Tuple[int, int, int]
    duration_ms: float
    scales: np.ndarray
    phi_deg: float


def request_stream(n: int, rate_hz: float = 20.0, seed: int = 4, lo: int = 32, hi: int = 128,
                   dur_ms=(200.0, 1000.0)) -> List[Request]:
    """Drug-screen requests with Poisson arrivals: each axis log-uniform in
    [lo, hi], duration uniform in dur_ms, scales uniform in [0, 1]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(n):
        t += rng.exponential(1.0 / rate_hz)
        dims = _log_uniform_dims(rng, (lo,) * 3, (hi,) * 3)
        out.append(Request(t, dims, float(rng.uniform(*dur_ms)), rng.uniform(0.0, 1.0, 3),
                           float(rng.uniform(-90.0, 90.0))))
    return out


# Niederer et al. (2011) N-version benchmark, as the reference's decks set it
# up (test/tests/NiedererTest/tools/object.data.proto, mkJobs.sh): a
# 3 x 7 x 20 mm brick of epicardial cells (cell type 102), fibres along the
# long (z) axis, the `fibre` conductivity with sigmaLi/sigmaTi below, a box
# stimulus of -35.71429 uA/uF for 2 ms in the corner cube x, y, z < stimBox.
NIEDERER_SIGMA_L = 0.0001334177
NIEDERER_SIGMA_T = 0.0000176062
NIEDERER_STIMBOX = {0.5: 4, 0.2: 8, 0.1: 16}
# published activation times (ms) at P1..P8 and the centre, Cardioid r145
# (test/tests/NiedererTest/tools/table.tex:5-27), keyed by (dx, dt)
NIEDERER_TABLE = {
    (0.2, 0.01): [1.24, 35.09, 11.15, 36.43, 39.83, 53.25, 41.58, 54.18, 24.05],
    (0.2, 0.005): [1.23, 34.72, 11.09, 36.05, 39.63, 52.73, 41.34, 53.65, 23.80],
    (0.5, 0.01): [1.24, 44.43, 20.79, 47.25, 107.97, 119.63, 109.29, 120.47, 55.38],
}
NIEDERER_POINTS = ["P1", "P2", "P3", "P4", "P5", "P6", "P7", "P8", "C"]


def niederer(dx: float = 0.2):
    """(anatomy, stimuli, point cell indices) of the Niederer slab at dx.
    The points follow the reference's getTimes.py: corners P1..P8 and the
    centre (nx//2, ny//2, nz//2)."""
    nx, ny, nz = int(3 / dx), int(7 / dx), int(20 / dx)
    n = nx * ny * nz
    f = np.tile(np.array([0.0, 0.0, 1.0]), (n, 1))
    a = Anatomy(nx, ny, nz, dx, dx, dx, np.arange(n, dtype=np.int64), np.zeros(n, np.uint8),
                fibre_tensor(f, NIEDERER_SIGMA_L, NIEDERER_SIGMA_T))
    sb = NIEDERER_STIMBOX[dx]
    stim = [box_stimulus(xMax=sb, yMax=sb, zMax=sb, nx=nx, ny=ny, nz=nz, vStim=-35.71429,
                         duration=2.0)]
    X, Y, Z = nx - 1, ny - 1, nz - 1
    corners = [(0, 0, 0), (0, 0, Z), (X, 0, 0), (X, 0, Z), (0, Y, 0), (0, Y, Z), (X, Y, 0),
               (X, Y, Z), (nx // 2, ny // 2, nz // 2)]
    idx = [x + nx * (y + ny * z) for x, y, z in corners]
    return a, stim, idx


@dataclass
class Share:
    """One part of an anatomy split along x: its own cells, and the remote
    cells (other parts' tissue inside its padded bounding box) whose Vm it
    reads from its neighbours each step."""
    anatomy: Anatomy
    remote_gid: np.ndarray
    remote_sigma: np.ndarray      # (nremote, 6)
    x_range: Tuple[int, int]


def split_x(anat: Anatomy, parts: int) -> List[Share]:
    """Cuts the anatomy into `parts` slabs of whole x planes with about equal
    cell counts (the slowest box axis, so each part's cells and box are
    contiguous): the layout one GPU of a `parts`-GPU run would hold."""
    x, y, z = anat.tuples()
    counts = np.bincount(x, minlength=anat.nx)
    cum = np.cumsum(counts)
    cuts = [0] + [int(np.searchsorted(cum, anat.ncell * p / parts, side="left")) + 1
                  for p in range(1, parts)] + [anat.nx]
    cuts = sorted(set(min(max(c, 0), anat.nx) for c in cuts))
    sig = anat.sigma.reshape(-1, 6)
    out = []
    for a, b in zip(cuts[:-1], cuts[1:]):
        own = (x >= a) & (x < b)
        if not own.any():
            continue
        lo = [x[own].min() - 1, y[own].min() - 1, z[own].min() - 1]
        hi = [x[own].max() + 1, y[own].max() + 1, z[own].max() + 1]
        near = (~own & (x >= lo[0]) & (x <= hi[0]) & (y >= lo[1]) & (y <= hi[1])
                & (z >= lo[2]) & (z <= hi[2]))
        part = Anatomy(anat.nx, anat.ny, anat.nz, anat.dx, anat.dy, anat.dz, anat.gid[own],
                       anat.celltype[own], sig[own])
        out.append(Share(part, anat.gid[near], sig[near], (a, b)))
    return out
