"""The two-scale decomposition walk's index arithmetic, restated on the CPU.

`k_iuwt_dec2` (pymoresane_amd/csrc/iuwt.hip) decomposes scales s and s+1 in one walk over the
comb/phase frame of dilation d = 2^s: a block owns one row phase, a chunk of Q column phases and TJ
teeth, stages C_s rows with 10 halo teeth in LDS, keeps C_{s+1} for 8 halo teeth on each side in a
second LDS row, and produces w_s, C_{s+2} and w_{s+1} at lags of 4, 7 and 11 rows behind its input.
This module restates the kernel's plan (`dec2_plan`), block decomposition, lane and LDS indices, step
tables, border folding and mirrored accumulation orders in numpy (float64 arithmetic), one block
after another and all lanes of a step at once, and checks on the CPU that

* every LDS read of a live lane lands on a staged slot (unstaged slots hold NaN here) and every
  global store offset lies inside the plane;
* every output pixel of the three planes is written exactly once;
* the planes equal the float64 oracle's w_s, w_{s+1} and C_{s+2} (to rounding), on square and
  non-square planes, odd sides, every dilation the kernel takes (1, 2, 4, 8) and the three walk
  lengths the autotuner's block targets give.

The GPU tests check the kernel itself: bit-identical to one launch per scale.
"""

import numpy as np
import pytest

from oracle import iuwt64

H5 = np.array([1, 4, 6, 4, 1], float) / 16
H9 = np.array([1, 8, 28, 56, 70, 56, 28, 8, 1], float) / 256

def fold_idx(i, n):
    i = np.where(i < 0, i + 2 * n, i)
    i = np.where(i >= 2 * n, i - 2 * n, i)
    return np.where(i >= n, 2 * n - 1 - i, i)

def fold_mirrors(i, n):
    i = np.where(i < 0, i + 2 * n, i)
    i = np.where(i >= 2 * n, i - 2 * n, i)
    return i >= n

def plan(h, w, s, target=2048):
    """dec2_plan: launch shape for scales (s, s+1), None where the kernel does not apply."""
    d = 1 << s
    if d > 8 or 20 * d > h or 20 * d > w:
        return None
    q = d
    nt = 256 if q <= 2 else 512
    teeth = -(-w // d)
    tj = min(nt // q - 16, teeth)
    if tj < 8:
        return None
    mr = -(-h // d)
    strips = d * (-(-d // q)) * (-(-teeth // tj))
    segs = max(1, -(-target // strips))
    k = max(-(-mr // segs), 40)
    k = min(k, 128 - 21, mr)
    return dict(d=d, q=q, nt=nt, tj=tj, k_rows=k, n_cg=-(-d // q), n_tb=-(-teeth // tj), n_seg=-(-mr // k),
                blocks=d * (-(-d // q)) * (-(-teeth // tj)) * (-(-mr // k)))

def run(cs, s, target=2048):
    """w_s, w_{s+1}, C_{s+2} from C_s as the kernel's blocks compute them (NaN where no block wrote)."""
    h, w = cs.shape
    pl = plan(h, w, s, target)
    assert pl is not None
    d, Q, NT, TJ = pl['d'], pl['q'], pl['nt'], pl['tj']
    ML = 2
    HT, HC, LAG = 8, 10, 10
    c_elems = (TJ + 2 * HC) * Q
    t_elems = (TJ + 2 * HT) * Q
    assert c_elems <= ML * NT and t_elems <= NT
    flat = cs.reshape(-1)
    w0 = np.full(h * w, np.nan); w1 = np.full(h * w, np.nan); cc = np.full(h * w, np.nan)
    hits = np.zeros(h * w, int)
    for b in range(pl['blocks']):
        seg = b % pl['n_seg']; bb = b // pl['n_seg']
        tb = bb % pl['n_tb']; bb //= pl['n_tb']
        cg = bb % pl['n_cg']; p = bb // pl['n_cg']
        t = np.arange(NT)
        q = t & (Q - 1)
        jt = (t >> int(np.log2(Q))) - HT
        t_live = t < t_elems
        o_live = (jt >= 0) & (jt < TJ)
        col = cg * Q + q + (tb * TJ + jt) * d
        col_ok = o_live & (col < w)
        tl_mir = t_live & fold_mirrors(col, w)
        c_lds = np.where(t_live, t + 2 * Q, 2 * Q)
        o_lds = np.where(o_live, t, HT * Q)
        mr = (h - p + d - 1) // d
        o_begin = seg * pl['k_rows']; o_end = min(o_begin + pl['k_rows'], mr)
        n_out = o_end - o_begin
        if n_out <= 0:
            continue
        i_begin = o_begin - LAG
        n_in = n_out + 2 * LAG
        n_steps = n_in + 1
        assert n_steps <= 128
        e = t[None, :] + NT * np.arange(ML)[:, None]
        jj = e >> int(np.log2(Q)); qq = e & (Q - 1)
        ccol = cg * Q + qq + (tb * TJ + jj - HC) * d
        assert np.all((ccol[e < c_elems] >= -2 * w) & (ccol[e < c_elems] < 4 * w))
        ld = np.where(e < c_elems, fold_idx(ccol, w), 0)
        va = np.zeros((5, NT)); vb = np.zeros((9, NT)); vx = np.zeros((5, NT))
        v2a = np.zeros((9, NT)); v2b = np.zeros((17, NT)); v2x = np.zeros((9, NT))
        tbuf = [np.zeros(NT), np.zeros(NT)]
        for st in range(n_steps):
            r_in = i_begin + min(st, n_in - 1)
            nat = p + r_in * d
            assert -2 * h <= nat < 4 * h
            row = flat[fold_idx(np.array(nat), h) * w:][:w]
            cbuf = np.zeros(ML * NT)
            cbuf[e.reshape(-1)] = row[ld.reshape(-1)]
            cbuf[(e >= c_elems).reshape(-1)] = np.nan  # slots no live lane may read
            # stage 1
            idx = c_lds[None, :] + (np.arange(5)[:, None] - 2) * Q
            assert idx.min() >= 0 and idx.max() < ML * NT
            x = cbuf[idx]
            ha = np.where(tl_mir, (H5[:, None] * x[::-1]).sum(0), (H5[:, None] * x).sum(0))
            va = np.roll(va, -1, 0); va[4] = ha
            mrow = p + (i_begin + st - 2) * d
            mir = fold_mirrors(np.array(mrow), h)
            c = (H5[:, None] * (va[::-1] if mir else va)).sum(0)
            tcur = tbuf[st & 1]; tprev = tbuf[(st + 1) & 1]
            idx = o_lds[None, :] + (np.arange(9)[:, None] - 2) * Q
            assert idx.min() >= 0 and idx.max() < ML * NT
            y = cbuf[idx]
            hb = (H9[:, None] * y).sum(0)
            vb = np.roll(vb, -1, 0); vb[8] = hb
            vx = np.roll(vx, -1, 0); vx[4] = y[4]
            wv = vx[0] - (H9[:, None] * vb).sum(0)
            orow = i_begin + st - 4
            if 0 <= st - 14 < n_out:
                m = col_ok
                off = (p + orow * d) * w + col[m]
                assert off.min() >= 0 and off.max() < h * w
                w0[off] = wv[m]; hits[off] += 1
            # stage 2 reads the previous step's row before this step's row replaces it in the other buffer
            idx = o_lds[None, :] + 2 * (np.arange(9)[:, None] - 4) * Q
            assert idx.min() >= 0 and idx.max() < NT and idx[:, o_live].max() < t_elems
            z = tprev[idx]
            ha2 = (H5[:, None] * z[2:7]).sum(0); hb2 = (H9[:, None] * z).sum(0)
            v2a = np.roll(v2a, -1, 0); v2a[8] = ha2
            v2b = np.roll(v2b, -1, 0); v2b[16] = hb2
            v2x = np.roll(v2x, -1, 0); v2x[8] = z[4]
            if 0 <= st - 17 < n_out:
                c2 = (H5[:, None] * v2a[0::2]).sum(0)
                off = (p + (i_begin + st - 7) * d) * w + col[col_ok]
                assert off.min() >= 0 and off.max() < h * w
                cc[off] = c2[col_ok]
            if 0 <= st - 21 < n_out:
                wv1 = v2x[0] - (H9[:, None] * v2b[0::2]).sum(0)
                off = (p + (i_begin + st - 11) * d) * w + col[col_ok]
                assert off.min() >= 0 and off.max() < h * w
                w1[off] = wv1[col_ok]
            tcur[:] = np.where(t_live, c, tcur)  # dead lanes write garbage in the kernel; keep it finite here
            tbuf[st & 1] = tcur
    assert np.all(hits == 1), (hits.min(), hits.max())
    return w0.reshape(h, w), w1.reshape(h, w), cc.reshape(h, w)

def check(h, w, s, seed=0, target=2048):
    """The restated walk against the oracle's decomposition of the same plane."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((h, w))
    det, sm = iuwt64.iuwt_decomposition(x, s + 2, store_smoothed=True)
    cs = x.copy()
    for k in range(s):
        cs = iuwt64.a_trous(cs, k)
    w0, w1, cc = run(cs, s, target)
    assert np.abs(w0 - det[s]).max() < 1e-9
    assert np.abs(w1 - det[s + 1]).max() < 1e-9
    assert np.abs(cc - sm).max() < 1e-9


@pytest.mark.parametrize("h,w,s,target", [
    (300, 260, 0, 2048), (257, 333, 0, 2048), (300, 260, 1, 4096), (200, 180, 2, 2048), (320, 336, 3, 2048),
    (160, 100, 0, 2048), (170, 161, 3, 1024), (1100, 400, 2, 4096), (160, 420, 3, 1024),
])
def test_two_scale_walk_indices_reproduce_the_oracle(h, w, s, target):
    check(h, w, s, target=target)


def test_two_scale_plan_rejects_what_the_kernel_does_not_take():
    assert plan(4096, 4096, 4) is None  # d = 16: chunks of 16 phases would idle half a block
    assert plan(100, 4096, 3) is None  # 20 d beyond a side
    assert plan(4096, 4096, 0)["k_rows"] <= 128 - 21
