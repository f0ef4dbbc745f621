"""The drop-in stencil's row exchange (cardioid_amd/csrc/row_shift.hpp),
modelled wavefront by wavefront on the CPU.

The kernel loads nine of the 19 neighbours and takes the ten x-1 / x+1
values from the neighbouring lanes, with the values that the wavefront's
end lanes (and lanes where the cell order leaves a row) need packed into
one more load: lane s of that load fetches value s of the list of
(needy lane, side, row) triples, and the needy lanes read theirs back by
lane permute.  This model follows the kernel's index arithmetic (ballots,
ranks, slots, the fallback to direct loads above 64 values) and checks
every value against a direct load, over sorted, ragged, sparse and
shuffled cell orders.  The GPU tests check the kernel itself bit for bit
against the oracle.
"""
import numpy as np
import pytest

# node deltas in NodeLocation order (fgr_weights.cpp) and the x rows
DELTA = [(0, 0, 0), (-1, 0, 0), (0, -1, 0), (1, 0, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1),
         (-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0), (-1, 0, -1), (0, -1, -1), (1, 0, -1),
         (0, 1, -1), (-1, 0, 1), (0, -1, 1), (1, 0, 1), (0, 1, 1)]
CENTRE, BACK, ON, LONE = [0, 2, 4, 5, 6], [1, 7, 10, 11, 15], [3, 8, 9, 13, 17], [12, 14, 16, 18]


def exchange(box, off, bl, act):
    """One wavefront: bl the lanes' box indices, act the lanes holding a
    cell.  Returns (values[64, 19], took the direct-load fallback)."""
    lanes = np.arange(64)
    v = np.full((64, 19), np.nan)
    for k in CENTRE + LONE:
        v[:, k] = box[bl + off[k]]
    bb = np.where(lanes > 0, np.roll(bl, 1), bl)            # from the previous lane
    bo = np.where(lanes < 63, np.roll(bl, -1), bl)          # from the next lane
    back_adj = (lanes > 0) & (bb + 1 == bl)
    on_adj = (lanes < 63) & (bo == bl + 1)
    MB, MO = act & ~back_adj, act & ~on_adj                 # the two ballots
    nb, no = int(MB.sum()), int(MO.sum())
    if 5 * (nb + no) > 64:
        for k in BACK + ON:
            v[:, k] = box[bl + off[k]]
        return v, True
    q, r = lanes // 5, lanes % 5
    sb = np.zeros(64, np.int64)
    side_on = np.zeros(64, bool)
    for j, c in enumerate(np.nonzero(MB)[0]):
        sb[q == j] = bl[c]
    for j, c in enumerate(np.nonzero(MO)[0], start=nb):
        sb[q == j] = bl[c]
        side_on[q == j] = True
    rel = np.where(side_on, off[np.array(ON)[r]], off[np.array(BACK)[r]])
    used = lanes < 5 * (nb + no)
    f = np.zeros(64)
    f[used] = box[sb[used] + rel[used]]
    below = lambda m: np.cumsum(m) - m                     # set bits below each lane
    src_b, src_o = 5 * below(MB), 5 * (nb + below(MO))
    for k in range(5):
        c = v[:, CENTRE[k]]
        up = np.where(lanes > 0, np.roll(c, 1), c)
        dn = np.where(lanes < 63, np.roll(c, -1), c)
        v[:, BACK[k]] = np.where(back_adj, up, f[(src_b + k) % 64])
        v[:, ON[k]] = np.where(on_adj, dn, f[(src_o + k) % 64])
    return v, False


@pytest.mark.parametrize("shape,frac,seed,order", [
    ((40, 30, 20), 0.6, 1, "gid"), ((64, 8, 8), 1.0, 2, "gid"), ((33, 17, 9), 0.97, 3, "gid"),
    ((20, 20, 20), 0.3, 4, "gid"), ((5, 5, 5), 1.0, 5, "gid"), ((70, 3, 3), 0.9, 6, "gid"),
    ((30, 20, 10), 0.8, 7, "shuffled"), ((14, 12, 10), 0.75, 8, "z_fastest")])
def test_every_exchanged_value_equals_a_direct_load(shape, frac, seed, order):
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    bx, by, bz = nx + 2, ny + 2, nz + 2
    off = np.array([dx + bx * (dy + by * dz) for dx, dy, dz in DELTA])
    for r in range(5):      # the host's check at create: rows along x
        assert off[BACK[r]] == off[CENTRE[r]] - 1 and off[ON[r]] == off[CENTRE[r]] + 1
    z, y, x = np.nonzero(rng.random((nz, ny, nx)) < frac)
    b = (x + 1) + bx * ((y + 1) + by * (z + 1))
    if order == "gid":
        b = np.sort(b)
    elif order == "shuffled":
        b = rng.permutation(b)
    else:
        b = b[np.lexsort((z, y, x))]
    box = rng.standard_normal(bx * by * bz)
    n, shared = b.size, 0
    for w0 in range(0, n, 64):
        i = w0 + np.arange(64)
        act = i < n
        bl = b[np.where(act, i, n - 1)]      # lanes past the last cell work on it
        v, direct = exchange(box, off, bl, act)
        shared += not direct
        want = box[bl[:, None] + off[None, :]]
        assert np.array_equal(v[act], want[act]), w0
    if order == "gid" and frac > 0.9 and nx >= 32:
        assert shared > 0       # long rows: the shared end load is exercised
