"""Set-up on the device: fibre conductivity (elec/FibreConductivity.cc:27-62)
and FGR weights (elec/FGRDiffusionOMP.cc:166-224) computed by HIP kernels.

Bar: bit for bit with the oracle (itself pinned against the reference's
FGRUtils.cc / FibreConductivity.cc and the mathematics, test_oracle_fgr.py)
for the weights and for integer fibre angles; within 2e-15 of sigma_L for
double angles, where the device's sin/cos stand in for glibc's.  The engine
built from device set-up runs bit for bit as one built from host set-up.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from cardioid_amd import hip, workloads as W

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


def _box(anat):
    """The grid padded by one on every side, z fastest (the reference's
    Array3d layout); tissue and sigma over it; the cells' box indices."""
    nx, ny, nz = anat.nx, anat.ny, anat.nz
    bx, by, bz = nx + 2, ny + 2, nz + 2
    g = anat.gid
    x, y, z = g % nx, (g // nx) % ny, g // (nx * ny)
    bidx = ((z + 1) + bz * ((y + 1) + by * (x + 1))).astype(np.int64)
    tissue = np.zeros(bx * by * bz, np.uint8)
    tissue[bidx] = 1
    sigma = np.zeros((bx * by * bz, 6))
    sigma[bidx] = anat.sigma.reshape(-1, 6)
    return (bx, by, bz), tissue, sigma, bidx


def _random_mask(n, frac, seed):
    """n^3 grid, each point tissue with probability frac, random fibres: every
    pattern of missing transverse nodes occurs"""
    rng = np.random.default_rng(seed)
    gid = np.nonzero(rng.uniform(size=n ** 3) < frac)[0].astype(np.int64)
    f = rng.normal(size=(gid.size, 3))
    f /= np.linalg.norm(f, axis=1, keepdims=True)
    return W.Anatomy(n, n, n, 0.2, 0.2, 0.2, gid, np.zeros(gid.size, np.uint8),
                     W.fibre_tensor(f))


@pytest.mark.parametrize("case", ["shell", "brick", "random_mask"])
def test_fgr_weights_device_equal_oracle(case):
    import oracle_lib as OL
    if case == "shell":
        anat, _ = W.ellipsoid_shell(48, 40, 64, seed=3)
    elif case == "brick":
        anat = W.brick(10, 9, 8)
    else:
        anat = _random_mask(24, 0.7, seed=4)
    say(f"device FGR weights vs oracle: {case}, {anat.ncell} cells")
    (bx, by, bz), tissue, sigma, bidx = _box(anat)
    n = bidx.size
    A = np.zeros(19 * n, np.float32)
    A0 = np.zeros(n)
    OL.oracle().ora_fgr_weights(bx, by, bz, OL.ptr(tissue, OL.u8p), OL.ptr(sigma), anat.dx,
                                anat.dy, anat.dz, n, OL.ptr(bidx, OL.i64p), OL.ptr(A, OL.fp),
                                OL.ptr(A0))
    dT = torch.from_numpy(tissue).cuda()
    dS = torch.from_numpy(sigma.reshape(-1)).cuda()
    dB = torch.from_numpy(bidx).cuda()
    dA0 = torch.zeros(n, dtype=torch.float64, device="cuda")
    dAw = torch.zeros(18 * n, dtype=torch.float32, device="cuda")
    hip.check(hip.lib().cdh_fgr_weights(bx, by, bz, anat.dx, anat.dy, anat.dz, vp(dT), vp(dS), n,
                                        vp(dB), vp(dA0), vp(dAw), n, None), "cdh_fgr_weights")
    torch.cuda.synchronize()
    got0 = dA0.cpu().numpy()
    got = dAw.cpu().numpy().reshape(18, n)
    want = A.reshape(n, 19)[:, 1:].T
    assert np.array_equal(got0, A0)
    assert np.array_equal(got.view(np.uint32), np.ascontiguousarray(want).view(np.uint32))
    assert (want != 0).sum() > n          # the case has real weights


def test_fibre_conductivity_int_equal_oracle():
    import oracle_lib as OL
    rng = np.random.default_rng(5)
    n = 50000
    th = rng.integers(-512, 513, n).astype(np.int32)
    ph = rng.integers(-512, 513, n).astype(np.int32)
    sl, st = 0.3e-3, 0.0315e-3
    dth, dph = torch.from_numpy(th).cuda(), torch.from_numpy(ph).cuda()
    out = torch.zeros(6 * n, dtype=torch.float64, device="cuda")
    hip.check(hip.lib().cdh_fibre_conductivity_int(n, vp(dth), vp(dph), sl, st, vp(out), None),
              "cdh_fibre_conductivity_int")
    got = out.cpu().numpy().reshape(n, 6)
    want = np.zeros((n, 6))
    row = np.zeros(6)
    for c in range(n):
        OL.oracle().ora_fibre_conductivity(th[c] * (np.pi / 256), ph[c] * (np.pi / 256), sl, st,
                                           OL.ptr(row))
        want[c] = row
    assert np.array_equal(got, want)
    bad = torch.tensor([0, 600], dtype=torch.int32, device="cuda")
    with pytest.raises(hip.CardioidHipError, match="outside"):
        hip.check(hip.lib().cdh_fibre_conductivity_int(2, vp(bad), vp(bad), sl, st, vp(out),
                                                       None), "cdh_fibre_conductivity_int")


def test_fibre_conductivity_double_near_oracle():
    import oracle_lib as OL
    rng = np.random.default_rng(6)
    n = 20000
    th, ph = rng.uniform(-4, 4, n), rng.uniform(-7, 7, n)
    sl, st = 0.1334e-3, 0.0176e-3
    out = torch.zeros(6 * n, dtype=torch.float64, device="cuda")
    dth, dph = torch.from_numpy(th).cuda(), torch.from_numpy(ph).cuda()
    hip.check(hip.lib().cdh_fibre_conductivity(n, vp(dth), vp(dph), sl, st, vp(out), None),
              "cdh_fibre_conductivity")
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(n, 6)
    row = np.zeros(6)
    err = 0.0
    for c in range(n):
        OL.oracle().ora_fibre_conductivity(th[c], ph[c], sl, st, OL.ptr(row))
        err = max(err, np.abs(got[c] - row).max())
    say(f"double angles: max |d sigma| {err:.3e} (sigma_L {sl})")
    assert err <= 2e-15 * sl


@pytest.mark.parametrize("mode", ["fp64_fused", "mixed_fused"])
def test_engine_device_setup_equals_host_setup(mode):
    a, depth = W.ellipsoid_shell(40, 36, 56, seed=2)
    stims = W.endocardial_point_stimuli(a, depth, 3, seed=2)
    say(f"engine set-up on device vs host: {mode}, {a.ncell} cells")
    out = []
    for host in (True, False):
        sim = hip.Simulation(a, W.CELLTYPES_EMI, 0.02, W.DIFFUSION_SCALE, stims, mode=mode,
                             setup_on_host=host)
        sim.run(400)
        out.append((sim.vm(), sim.state(), sim.activation()[0]))
        sim.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2])
    assert (out[0][2] >= 0).any()
