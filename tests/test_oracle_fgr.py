"""Pins for the oracle's FGR diffusion (CPU only).

The reference's FGRDiffusionOMP.cc cannot be built here (DiffusionUtils.cc
needs <mpi.h>), and its diffusionTest golden needs an anatomy blob missing
from the snapshot.  So the gradient stencils are compared bitwise with the
reference's FGRUtils.cc (compiled into oracle/_ref), and the assembled
operator is pinned by what the mathematics fixes:
  * isotropic uniform conductivity reduces it to the 7-point Laplacian;
  * it is exact for quadratic fields in the interior (so every diagonal and
    cross-derivative weight has the right sign and index);
  * it conserves charge on masked anatomies (flux out of one cell enters
    its neighbour), up to the float rounding of the weights;
  * rows sum to zero.
"""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import fp, i64p, ip, oracle, ptr, ref, ref_available, u8p

needs_ref = pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built")


@needs_ref
def test_gradient_stencils_bitwise_against_reference():
    O, R = oracle(), ref()
    rng = np.random.default_rng(0)
    pats = [np.ones(19, np.int32), np.zeros(19, np.int32)]
    pats += [rng.integers(0, 2, 19).astype(np.int32) for _ in range(6000)]
    for t in pats:
        t = np.ascontiguousarray(t)
        for face in range(6):
            a = np.zeros(57)
            b = np.zeros(57)
            R.ref_fgr_f2(face, ptr(t, ip), ptr(a))
            O.ora_fgr_f2(face, ptr(t, ip), ptr(b))
            assert np.array_equal(a, b), (face, t)


def build(tissue3, sigma6, h=(0.2, 0.2, 0.2)):
    """tissue3: (X,Y,Z) 0/1 mask without padding; sigma6: (X,Y,Z,6).
    Returns the padded box, block indices of tissue cells and weights."""
    X, Y, Z = tissue3.shape
    bx, by, bz = X + 2, Y + 2, Z + 2
    tis = np.zeros((bx, by, bz), np.uint8)
    tis[1:-1, 1:-1, 1:-1] = tissue3
    sig = np.zeros((bx, by, bz, 6))
    sig[1:-1, 1:-1, 1:-1] = sigma6
    cells = np.argwhere(tis.astype(bool))          # x-major, z fastest = box order
    bi = (cells[:, 2] + bz * (cells[:, 1] + by * cells[:, 0])).astype(np.int64)
    n = bi.size
    A = np.zeros(19 * n, np.float32)
    A0 = np.zeros(n)
    tisf = np.ascontiguousarray(tis.ravel())
    sigf = np.ascontiguousarray(sig.reshape(-1))
    oracle().ora_fgr_weights(bx, by, bz, ptr(tisf, u8p), ptr(sigf), h[0], h[1], h[2], n,
                             ptr(bi, i64p), ptr(A, fp), ptr(A0))
    off = np.zeros(19, np.int64)
    oracle().ora_fgr_offsets(bx, by, bz, ptr(off, i64p))
    return dict(shape=(bx, by, bz), cells=cells, bi=bi, A=A.reshape(n, 19), A0=A0, off=off,
                Af=A)


def apply(B, phi_box, scale=1.0):
    n = B["bi"].size
    out = np.zeros(n)
    phi = np.ascontiguousarray(phi_box.ravel(), dtype=np.float64)
    oracle().ora_fgr_calc(n, ptr(B["bi"], i64p), ptr(B["Af"], fp), ptr(B["A0"]),
                          ptr(B["off"], i64p), ptr(phi), scale, ptr(out))
    return out


def spd_tensor(rng, n):
    """random SPD tensors as (a11,a12,a13,a22,a23,a33)"""
    M = rng.normal(size=(n, 3, 3)) * 0.3
    S = M @ M.transpose(0, 2, 1) + np.eye(3) * 0.5
    return np.stack([S[:, 0, 0], S[:, 0, 1], S[:, 0, 2], S[:, 1, 1], S[:, 1, 2], S[:, 2, 2]], -1)


def test_isotropic_reduces_to_7_point_laplacian():
    sig = 0.1334e-3
    t = np.ones((6, 5, 7), np.uint8)
    s6 = np.zeros((6, 5, 7, 6))
    s6[..., 0] = s6[..., 3] = s6[..., 5] = sig
    h = (0.2, 0.25, 0.1)
    B = build(t, s6, h)
    A = B["A"]
    # NodeLocation: ZZZ, MZZ, ZMZ, PZZ, ZPZ, ZZM, ZZP, then 12 edge nodes
    interior = np.all((B["cells"] > 1) & (B["cells"] < np.array(B["shape"]) - 2), axis=1)
    for k, hk in ((1, h[0]), (3, h[0]), (2, h[1]), (4, h[1]), (5, h[2]), (6, h[2])):
        np.testing.assert_allclose(A[interior, k], sig / hk ** 2, rtol=1e-6)
    assert np.all(A[:, 7:] == 0)
    assert np.all(B["A0"] < 0)


@pytest.mark.parametrize("field", ["xx", "yy", "zz", "xy", "xz", "yz", "lin"])
def test_exact_for_quadratics_in_interior(field):
    rng = np.random.default_rng(1)
    X, Y, Z = 8, 7, 9
    s = spd_tensor(rng, 1)[0] * 1e-3
    t = np.ones((X, Y, Z), np.uint8)
    s6 = np.broadcast_to(s, (X, Y, Z, 6)).copy()
    h = (0.2, 0.15, 0.25)
    B = build(t, s6, h)
    bx, by, bz = B["shape"]
    gx, gy, gz = np.meshgrid(np.arange(bx) * h[0], np.arange(by) * h[1], np.arange(bz) * h[2],
                             indexing="ij")
    a11, a12, a13, a22, a23, a33 = s
    phi, expect = {
        "xx": (gx ** 2, 2 * a11), "yy": (gy ** 2, 2 * a22), "zz": (gz ** 2, 2 * a33),
        "xy": (gx * gy, 2 * a12), "xz": (gx * gz, 2 * a13), "yz": (gy * gz, 2 * a23),
        "lin": (0.3 * gx - 1.1 * gy + 0.7 * gz, 0.0),
    }[field]
    out = apply(B, phi)
    inner = np.all((B["cells"] > 2) & (B["cells"] < np.array(B["shape"]) - 3), axis=1)
    scale = max(abs(s).max(), 1e-30)
    # float weights: relative error of order 1e-7 of |A|*|phi|
    tol = 1e-6 * scale * (np.abs(phi).max() / min(h) ** 2)
    np.testing.assert_allclose(out[inner], expect, atol=tol)


def test_conservation_on_masked_anatomy():
    rng = np.random.default_rng(2)
    X, Y, Z = 14, 12, 10
    t = (rng.random((X, Y, Z)) < 0.7).astype(np.uint8)
    s6 = spd_tensor(rng, X * Y * Z).reshape(X, Y, Z, 6) * 1e-3
    B = build(t, s6)
    bx, by, bz = B["shape"]
    phi = np.zeros((bx, by, bz))
    c = B["cells"]
    phi[c[:, 0], c[:, 1], c[:, 2]] = rng.uniform(-85, 30, len(c))
    out = apply(B, phi)
    assert abs(out.sum()) < 1e-6 * np.abs(out).sum()
    # rows sum to zero: the centre weight is minus the sum of the others
    np.testing.assert_allclose(B["A0"], -B["A"][:, 1:].astype(np.float64).sum(1), rtol=1e-12,
                               atol=1e-300)
    const = apply(B, np.where(phi != 0, 7.0, 0.0) + 0 * phi)
    assert np.abs(const).max() < 1e-12 * np.abs(B["A0"]).max() * 7.0 * 19


def test_offsets_follow_array3d_layout():
    off = np.zeros(19, np.int64)
    oracle().ora_fgr_offsets(5, 6, 7, ptr(off, i64p))
    # z fastest: index = z + nz*(y + ny*x) (Array3d.hh:173-177)
    dz, dy, dx = 1, 7, 42
    expect = [0, -dx, -dy, dx, dy, -dz, dz, -dx - dy, dx - dy, dx + dy, -dx + dy,
              -dx - dz, -dy - dz, dx - dz, dy - dz, -dx + dz, -dy + dz, dx + dz, dy + dz]
    assert off.tolist() == expect


def _fibre(L, fn, theta, phi, sl, st):
    out = np.zeros(6)
    getattr(L, fn)(theta, phi, sl, st, ptr(out))
    return out


@needs_ref
def test_fibre_conductivity_bitwise_against_reference():
    """ora_fibre_conductivity equals the reference's FibreConductivity::compute
    (elec/FibreConductivity.cc, built into oracle/_ref) bit for bit, for the
    integer-angle grid of pio anatomies and for random double angles."""
    O, R = oracle(), ref()
    sl, st = 0.3e-3, 0.0315e-3
    rng = np.random.default_rng(1)
    angles = [(i * (np.pi / 256), j * (np.pi / 256)) for i in range(-256, 257, 7)
              for j in range(-512, 513, 11)]
    angles += [tuple(rng.uniform(-7, 7, 2)) for _ in range(3000)]
    for th, ph in angles:
        a = _fibre(R, "ref_fibre_conductivity", th, ph, sl, st)
        b = _fibre(O, "ora_fibre_conductivity", th, ph, sl, st)
        assert np.array_equal(a, b), (th, ph, a, b)


def test_fibre_conductivity_closed_form():
    """sigma = sT I + (sL - sT) f f^T with f = (sin t cos p, sin t sin p, cos t):
    the fibre direction has eigenvalue sL, its normals sT (so a dropped term,
    a swapped angle or a wrong sign fails)."""
    O = oracle()
    sl, st = 0.1334e-3, 0.0176e-3
    rng = np.random.default_rng(2)
    for _ in range(2000):
        th, ph = rng.uniform(-4, 4, 2)
        s = _fibre(O, "ora_fibre_conductivity", th, ph, sl, st)
        f = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
        M = st * np.eye(3) + (sl - st) * np.outer(f, f)
        want = np.array([M[0, 0], M[0, 1], M[0, 2], M[1, 1], M[1, 2], M[2, 2]])
        np.testing.assert_allclose(s, want, rtol=0, atol=2e-15 * sl)
    # theta = 0: fibres along z
    s = _fibre(O, "ora_fibre_conductivity", 0.0, 0.3, sl, st)
    np.testing.assert_allclose(s, [st, 0, 0, st, 0, sl], rtol=0, atol=1e-20)
