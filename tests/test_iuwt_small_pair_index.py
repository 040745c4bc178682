"""Index check of the two-scale small-plane kernels (k_iuwt_small2_dec / k_iuwt_small2_rec in
csrc/iuwt.hip), on the CPU.

This module restates the kernels' tile grid, staged-region and slot indices, and the tap order of
every chain (reversed at mirrored slot positions), and checks them against the one-scale kernel's
taps at the folded positions (fold_idx of csrc/pms_common.h): each output pixel is written by one
thread, every LDS index is inside its array, and every chain visits the plane positions that the
one-scale chain visits, in the same order.  The kernels form each 2-D sum separably (row sums of a
whole region first, then column sums over them), which visits the same positions in the same order
as the direct chains restated here.  The GPU tests then check the planes bit-for-bit.
"""

import numpy as np
import pytest

T = 16  # tile side in lattice steps


def ext(u, n):
    """Folded index and mirror flag of any integer lattice coordinate on the symmetric extension."""
    m = np.mod(u, 2 * n)
    mir = m >= n
    return np.where(mir, 2 * n - 1 - m, m), mir


def fold(i, n):
    """fold_idx: valid for i in [-2n, 3n)."""
    i = np.asarray(i)
    assert np.all(i >= -2 * n) and np.all(i < 3 * n)
    i = np.where(i < 0, i + 2 * n, i)
    i = np.where(i >= 2 * n, i - 2 * n, i)
    return np.where(i >= n, 2 * n - 1 - i, i)


def over_long(h, w, s):
    return (2 << s) > h or (2 << s) > w


def tiles(h, w, s):
    d = 1 << s
    ntr = -(-(-(-h // d)) // T)
    ntc = -(-(-(-w // d)) // T)
    for by in range(d * ntr):
        for bx in range(d * ntc):
            yield d, by // ntr, (by % ntr) * T, bx // ntc, (bx % ntc) * T


CASES = [((40, 40), 0), ((40, 40), 2), ((17, 23), 1), ((64, 48), 3), ((256, 256), 6), ((255, 201), 4),
         ((200, 31), 1), ((128, 512), 5)]


@pytest.mark.parametrize("shape,s", CASES)
def test_small_pair_decomposition_indices(shape, s):
    h, w = shape
    assert not over_long(h, w, s + 1)
    written = np.zeros(shape, np.int32)
    ty, tx = np.divmod(np.arange(256), 16)
    u, v = np.divmod(np.arange(32 * 32), 32)
    for d, pr, i0, pc, j0 in tiles(h, w, s):
        ry, my = ext(pr + d * (i0 - 10 + np.arange(36)), h)
        rx, mx = ext(pc + d * (j0 - 10 + np.arange(36)), w)
        # C_{s+1} slot (u, v) sits at staged (u + 2, v + 2); its chains visit staged rows u + i and
        # columns v + b, reversed where the slot row or column is mirrored
        P_r, P_c = ry[u + 2], rx[v + 2]
        for i in range(5):
            row = np.where(my[u + 2], u + 4 - i, u + i)
            assert row.max() < 36
            np.testing.assert_array_equal(ry[row], fold(P_r + (i - 2) * d, h))
        for b in range(5):
            col = np.where(mx[v + 2], v + 4 - b, v + b)
            assert col.max() < 36
            np.testing.assert_array_equal(rx[col], fold(P_c + (b - 2) * d, w))
        slot_r, slot_c = P_r.reshape(32, 32)[:, 0], P_c.reshape(32, 32)[0]
        r = pr + d * (i0 + ty)
        c = pc + d * (j0 + tx)
        live = (r < h) & (c < w)
        np.add.at(written, (r[live], c[live]), 1)
        rl, cl, tyl, txl = r[live], c[live], ty[live], tx[live]
        # w_s: staged (ty + 6 + i, tx + 6 + b), centre (ty + 10, tx + 10)
        np.testing.assert_array_equal(ry[tyl + 10], rl)
        np.testing.assert_array_equal(rx[txl + 10], cl)
        for i in range(9):
            np.testing.assert_array_equal(ry[tyl + 6 + i], fold(rl + (i - 4) * d, h))
            np.testing.assert_array_equal(rx[txl + 6 + i], fold(cl + (i - 4) * d, w))
        # w_{s+1}: slots (ty + 2i, tx + 2b), centre (ty + 8, tx + 8); C_{s+2}: (ty + 4 + 2k, tx + 4 + 2b)
        np.testing.assert_array_equal(slot_r[tyl + 8], rl)
        np.testing.assert_array_equal(slot_c[txl + 8], cl)
        for i in range(9):
            assert (tyl + 2 * i).max() < 32
            np.testing.assert_array_equal(slot_r[tyl + 2 * i], fold(rl + (i - 4) * 2 * d, h))
            np.testing.assert_array_equal(slot_c[txl + 2 * i], fold(cl + (i - 4) * 2 * d, w))
        for k in range(5):
            np.testing.assert_array_equal(slot_r[tyl + 4 + 2 * k], fold(rl + (k - 2) * 2 * d, h))
            np.testing.assert_array_equal(slot_c[txl + 4 + 2 * k], fold(cl + (k - 2) * 2 * d, w))
    np.testing.assert_array_equal(written, 1)


@pytest.mark.parametrize("shape,s", CASES)
def test_small_pair_recomposition_indices(shape, s):
    h, w = shape
    assert not over_long(h, w, s + 1)
    written = np.zeros(shape, np.int32)
    ty, tx = np.divmod(np.arange(256), 16)
    n_slots = 20 * 20
    e = np.minimum(np.arange(2 * 256), n_slots - 1)  # clamped: repeated slots write the same value
    assert set(e.tolist()) == set(range(n_slots))
    u, v = np.divmod(e, 20)
    for d, pr, i0, pc, j0 in tiles(h, w, s):
        ry, my = ext(pr + d * (i0 - 6 + np.arange(28)), h)
        rx, mx = ext(pc + d * (j0 - 6 + np.arange(28)), w)
        # T slot (u, v) sits at staged (u + 4, v + 4); its chains visit staged rows u + 2i and
        # columns v + 2b (spacing 2 on the lattice: scale s+1), reversed where mirrored
        P_r, P_c = ry[u + 4], rx[v + 4]
        for i in range(5):
            row = np.where(my[u + 4], u + 8 - 2 * i, u + 2 * i)
            assert row.max() < 28
            np.testing.assert_array_equal(ry[row], fold(P_r + (i - 2) * 2 * d, h))
        for b in range(5):
            col = np.where(mx[v + 4], v + 8 - 2 * b, v + 2 * b)
            assert col.max() < 28
            np.testing.assert_array_equal(rx[col], fold(P_c + (b - 2) * 2 * d, w))
        slot_r, slot_c = ry[4:24], rx[4:24]
        r = pr + d * (i0 + ty)
        c = pc + d * (j0 + tx)
        live = (r < h) & (c < w)
        np.add.at(written, (r[live], c[live]), 1)
        rl, cl, tyl, txl = r[live], c[live], ty[live], tx[live]
        # R_s: T slots (ty + i, tx + b) at spacing 1
        for i in range(5):
            assert (tyl + i).max() < 20
            np.testing.assert_array_equal(slot_r[tyl + i], fold(rl + (i - 2) * d, h))
            np.testing.assert_array_equal(slot_c[txl + i], fold(cl + (i - 2) * d, w))
    np.testing.assert_array_equal(written, 1)


def test_mirror_flag_reverses_taps():
    """A position on a mirrored period sees the folded position's taps in reverse order, one on an
    unmirrored period (two reflections) in the same order -- the case analysis the kernels rely on."""
    n, d = 23, 2
    for u in range(-3 * n, 4 * n):
        p, mir = ext(u, n)
        taps_u = [ext(u + (k - 2) * d, n)[0] for k in range(5)]
        taps_p = [ext(p + (k - 2) * d, n)[0] for k in range(5)]
        assert taps_u == (taps_p[::-1] if mir else taps_p)
