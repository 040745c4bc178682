"""Pins of the float64 oracle against library routines, closed forms and brute force.

None of these retypes the oracle's formula: each compares it against an
independent computation (scipy routines, direct sums, invariants) chosen so
that a dropped tap, a wrong reflection, a transposed axis, a sign error or a
wrong crop offset fails at least one test.
"""

import numpy as np
import pytest
from scipy import ndimage, signal

from oracle import convolution64 as conv
from oracle import iuwt64 as iuwt
from oracle import toolbox64 as tools

RNG = np.random.default_rng(1234)


def dilated_b3(scale):
    d = 1 << scale
    k = np.zeros(4 * d + 1)
    k[::d] = iuwt.B3
    return k


@pytest.mark.parametrize("shape", [(64, 64), (40, 40), (48, 80), (32, 16)])
def test_a_trous_equals_scipy_correlate_reflect(shape):
    x = RNG.standard_normal(shape)
    s_max = int(np.log2(min(shape))) - 1
    for s in range(s_max + 1):
        ref = ndimage.correlate1d(x, dilated_b3(s), axis=0, mode="reflect")
        ref = ndimage.correlate1d(ref, dilated_b3(s), axis=1, mode="reflect")
        np.testing.assert_allclose(iuwt.a_trous(x, s), ref, rtol=0, atol=1e-13)


def test_a_trous_brute_force_direct_sum():
    """2-D direct sum over the explicit half-sample symmetric extension."""
    x = RNG.standard_normal((8, 12))
    h, w = x.shape
    ext = np.pad(x, ((h, h), (w, w)), mode="symmetric")  # d c b a | a b c d
    for s in range(2):
        d = 1 << s
        out = np.zeros_like(x)
        for i in range(h):
            for j in range(w):
                acc = 0.0
                for a in range(-2, 3):
                    for b in range(-2, 3):
                        acc += iuwt.B3[a + 2] * iuwt.B3[b + 2] * ext[h + i + a * d, w + j + b * d]
                out[i, j] = acc
        np.testing.assert_allclose(iuwt.a_trous(x, s), out, atol=1e-14)


def slice_a_trous(c0, scale):
    """The reference's slice arithmetic for one smoothing (iuwt.py:163-189): per tap an in-range
    shifted slice plus a reversed border slice, with numpy clipping both when the tap offset
    exceeds the axis; axis 0 first, then axis 1 (as axis 0 of the transpose)."""
    def axis0(a):
        out = iuwt.B3[2] * a
        for k, off in ((0, 2 << scale), (1, 1 << scale)):
            out[off:] += iuwt.B3[k] * a[:-off]
            out[:off] += iuwt.B3[k] * a[off - 1::-1]
        for k, off in ((3, 1 << scale), (4, 2 << scale)):
            out[:-off] += iuwt.B3[k] * a[off:]
            out[-off:] += iuwt.B3[k] * a[:-off - 1:-1]
        return out
    return axis0(axis0(np.asarray(c0, dtype=np.float64)).T.copy()).T


@pytest.mark.parametrize("shape", [(8, 8), (16, 12), (6, 10), (32, 32)])
def test_a_trous_equals_reference_slices_at_every_scale(shape):
    """Including scales whose taps are longer than the side (2**(s+1) > side), where numpy's
    clipped slices read the whole axis reversed."""
    x = RNG.standard_normal(shape)
    for s in range(int(np.log2(max(shape))) + 2):
        np.testing.assert_allclose(iuwt.a_trous(x, s), slice_a_trous(x, s), rtol=0, atol=1e-14)


def test_over_long_scales_keep_perfect_reconstruction():
    x = RNG.standard_normal((16, 16))
    det, smooth = iuwt.iuwt_decomposition(x, 6, store_smoothed=True)  # scales 3..5 exceed the side
    np.testing.assert_allclose(iuwt.iuwt_recomposition(det, 0, smooth), x, atol=1e-12)


@pytest.mark.parametrize("adjust", [0, 1, 2])
def test_iuwt_perfect_reconstruction(adjust):
    x = RNG.standard_normal((64, 64))
    det, smooth = iuwt.iuwt_decomposition(x, 5, store_smoothed=True)
    np.testing.assert_allclose(iuwt.iuwt_recomposition(det, 0, smooth), x, atol=1e-12)
    # with scale_adjust the coarse planes plus the pre-smoothed base reproduce the smoothed input
    det_a, smooth_a = iuwt.iuwt_decomposition(x, 5, scale_adjust=adjust, store_smoothed=True)
    base = x
    for s in range(adjust):
        base = iuwt.a_trous(base, s)
    rec = iuwt.iuwt_recomposition(det_a, adjust, smooth_a)
    expect = base
    for s in range(adjust - 1, -1, -1):
        expect = iuwt.a_trous(expect, s)
    np.testing.assert_allclose(rec, expect, atol=1e-12)
    np.testing.assert_allclose(det_a, det[adjust:], atol=1e-12)


def test_iuwt_constant_has_zero_details_and_is_linear():
    c = np.full((32, 32), 3.25)
    det, smooth = iuwt.iuwt_decomposition(c, 4, store_smoothed=True)
    assert np.max(np.abs(det)) < 1e-13
    np.testing.assert_allclose(smooth, c, atol=1e-13)
    a, b = RNG.standard_normal((2, 32, 32))
    np.testing.assert_allclose(iuwt.iuwt_decomposition(2 * a - 3 * b, 4),
                               2 * iuwt.iuwt_decomposition(a, 4) - 3 * iuwt.iuwt_decomposition(b, 4),
                               atol=1e-12)


def test_iuwt_delta_response_is_b3_autocorrelation():
    """A centred delta: scale-0 smoothing equals the outer product of B3, and the
    detail is delta minus the outer product of B3*B3 (away from the borders)."""
    x = np.zeros((32, 32))
    x[16, 16] = 1.0
    sm = iuwt.a_trous(x, 0)
    expect = np.zeros_like(x)
    expect[14:19, 14:19] = np.outer(iuwt.B3, iuwt.B3)
    np.testing.assert_allclose(sm, expect, atol=1e-15)
    det = iuwt.iuwt_decomposition(x, 1)[0]
    b9 = np.convolve(iuwt.B3, iuwt.B3)
    expect = -np.outer(b9, b9)
    expect[4, 4] += 1.0
    np.testing.assert_allclose(det[12:21, 12:21], expect, atol=1e-15)


@pytest.mark.parametrize("n", [8, 16, 32])
def test_fft_convolve_linear_equals_fftconvolve_crop(n):
    x = RNG.standard_normal((n, n))
    psf = RNG.standard_normal((2 * n, 2 * n))
    got = conv.fft_convolve(x, conv.psf_spectrum(psf), "linear")
    ref = signal.fftconvolve(x, psf)[n:2 * n, n:2 * n]
    np.testing.assert_allclose(got, ref, atol=1e-11)


def test_fft_convolve_linear_brute_force():
    n = 6
    x = RNG.standard_normal((n, n))
    psf = RNG.standard_normal((2 * n, 2 * n))
    out = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            for u in range(n):
                for v in range(n):
                    out[i, j] += x[u, v] * psf[i - u + n, j - v + n]
    np.testing.assert_allclose(conv.fft_convolve(x, conv.psf_spectrum(psf), "linear"), out, atol=1e-12)


def test_fft_convolve_linear_same_size_psf_is_padded():
    """main.py:203-206: an N x N PSF is zero-padded to 2N; its centre N/2 lands on N."""
    n = 16
    x = RNG.standard_normal((n, n))
    psf = RNG.standard_normal((n, n))
    got = conv.fft_convolve(x, conv.psf_spectrum(conv.pad_array(psf)), "linear")
    big = np.zeros((2 * n, 2 * n))
    big[n // 2:n // 2 + n, n // 2:n // 2 + n] = psf
    ref = signal.fftconvolve(x, big)[n:2 * n, n:2 * n]
    np.testing.assert_allclose(got, ref, atol=1e-11)


def test_fft_convolve_circular_brute_force():
    n = 8
    x = RNG.standard_normal((n, n))
    psf = RNG.standard_normal((n, n))
    out = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            for u in range(n):
                for v in range(n):
                    out[i, j] += x[u, v] * psf[(i - u + n // 2) % n, (j - v + n // 2) % n]
    np.testing.assert_allclose(conv.fft_convolve(x, conv.psf_spectrum(psf), "circular"), out, atol=1e-12)


def test_estimate_threshold_known_medians():
    # odd count: a known middle value; even count: the mean of the two middle values
    a = np.array([[-5.0, 1.0, 2.0], [-0.5, 7.0, 3.0], [4.0, -9.0, 0.25]])  # |.| sorted: .25 .5 1 2 3 4 5 7 9
    b = np.array([[1.0, -2.0], [3.0, -10.0]])
    assert tools.estimate_threshold(a[None])[0] == pytest.approx(3.0 / 0.6745, rel=1e-15)
    assert tools.estimate_threshold(b[None])[0] == pytest.approx(2.5 / 0.6745, rel=1e-15)


def test_estimate_threshold_masks():
    x = np.arange(1.0, 101.0).reshape(1, 10, 10)
    # edge_excl=2 keeps rows/cols 2..7 -> values r*10+c+1 for r,c in 2..7: median of those 36
    inner = x[0, 2:8, 2:8].ravel()
    assert tools.estimate_threshold(x, edge_excl=2)[0] == pytest.approx(np.sort(inner)[17:19].mean() / 0.6745)
    # int_excl=1 removes the 2x2 square at rows/cols 4..5 (mid = 5)
    keep = np.ones((10, 10), bool)
    keep[4:6, 4:6] = False
    vals = np.sort(x[0][keep])
    assert tools.estimate_threshold(x, int_excl=1)[0] == pytest.approx(vals[47:49].mean() / 0.6745)


def test_estimate_threshold_int_excl_uses_the_row_count_for_both_axes():
    """iuwt_toolbox.py:29,40: mid = in1.shape[1]/2 -- the row count of the (scales, rows, cols)
    stack -- bounds the excluded square on both axes, so a 6 x 14 plane with int_excl=2 drops rows
    and columns 1..4 (mid 3), not columns 5..8 around the column centre."""
    x = np.arange(1.0, 85.0).reshape(1, 6, 14)
    keep = np.ones((6, 14), bool)
    keep[1:5, 1:5] = False
    vals = np.sort(x[0][keep])  # 68 values: the median is the mean of ranks 33 and 34
    assert tools.estimate_threshold(x, int_excl=2)[0] == pytest.approx(vals[33:35].mean() / 0.6745, rel=1e-15)


def test_apply_threshold_is_strict_and_keeps_negatives():
    w = np.array([[[4.0, -4.0, 4.0000001, -4.5, 0.1]]])
    out = tools.apply_threshold(w, np.array([1.0]), sigma_level=4)
    np.testing.assert_array_equal(out, [[[0.0, 0.0, 4.0000001, -4.5, 0.0]]])


def bfs_labels(fg):
    """Brute-force 8-connected labelling in raster order of the first pixel."""
    h, w = fg.shape
    lab = np.zeros((h, w), dtype=np.int32)
    nxt = 0
    for i in range(h):
        for j in range(w):
            if fg[i, j] and lab[i, j] == 0:
                nxt += 1
                stack = [(i, j)]
                lab[i, j] = nxt
                while stack:
                    a, b = stack.pop()
                    for da in (-1, 0, 1):
                        for db in (-1, 0, 1):
                            u, v = a + da, b + db
                            if 0 <= u < h and 0 <= v < w and fg[u, v] and lab[u, v] == 0:
                                lab[u, v] = nxt
                                stack.append((u, v))
    return lab


@pytest.mark.parametrize("seed", range(4))
def test_labels_match_brute_force(seed):
    rng = np.random.default_rng(seed)
    plane = np.where(rng.random((24, 31)) < 0.35, rng.standard_normal((24, 31)), 0.0)
    lab, _ = tools.label_scale(plane)
    np.testing.assert_array_equal(lab, bfs_labels(plane != 0))


def test_source_extraction_nested_keep_rule():
    w = np.zeros((2, 12, 12))
    # coarse scale: two objects; only the left one reaches 0.75 * max
    w[1, 2:5, 1:4] = 1.0
    w[1, 8:10, 8:10] = 0.5
    # fine scale: object A overlaps the kept coarse object and is strong; object B is strong but
    # lies outside the coarse mask; object C (negative pixels joined to a positive one) overlaps
    w[0, 3, 2] = 2.0           # A
    w[0, 9, 9] = 2.0           # B (over the dropped coarse object)
    w[0, 2, 1] = -0.3          # C touches A diagonally? no: (2,1) is 8-adjacent to (3,2) -> same object
    w[0, 6, 6] = 1.0           # D: weak and outside -> dropped
    ext, mask = tools.source_extraction(w, 0.75)
    expect = np.zeros((2, 12, 12), dtype=np.int32)
    expect[1, 2:5, 1:4] = 1
    expect[0, 3, 2] = 1
    expect[0, 2, 1] = 1
    np.testing.assert_array_equal(mask, expect)
    np.testing.assert_array_equal(ext, expect * w)


def test_source_extraction_neg_comp_uses_abs():
    w = np.zeros((1, 6, 6))
    w[0, 1, 1] = -3.0
    w[0, 4, 4] = 1.0
    _, m_pos = tools.source_extraction(w, 0.75)
    _, m_abs = tools.source_extraction(w, 0.75, neg_comp=True)
    assert m_pos[0, 4, 4] == 1 and m_pos[0, 1, 1] == 0
    assert m_abs[0, 1, 1] == 1 and m_abs[0, 4, 4] == 0


def test_snr_ratio_closed_form():
    a = np.array([3.0, 4.0])      # |a| = 5
    b = np.array([3.0, 3.5])      # |a-b| = 0.5
    assert tools.snr_ratio(a, b) == pytest.approx(20.0)


def test_fast_backends_match_restatement():
    """The scipy a-trous backend and threaded transforms used for long oracle runs agree with the
    restatement / numpy.fft to rounding level."""
    from oracle import convolution64 as conv
    x = np.random.default_rng(9).standard_normal((48, 80))
    ref = [iuwt.a_trous(x, s) for s in range(4)]
    iuwt.use_scipy_backend(True)
    try:
        for s in range(4):
            assert np.max(np.abs(iuwt.a_trous(x, s) - ref[s])) < 1e-14
    finally:
        iuwt.use_scipy_backend(False)
    img = np.random.default_rng(10).standard_normal((32, 32))
    psf = np.random.default_rng(11).standard_normal((64, 64))
    a = conv.fft_convolve(img, conv.psf_spectrum(psf), "linear")
    conv.set_fft_workers(4)
    try:
        b = conv.fft_convolve(img, conv.psf_spectrum(psf), "linear")
    finally:
        conv.set_fft_workers(1)
    assert np.max(np.abs(a - b)) < 1e-12


def test_c1_crop_interior_matches_full_image():
    """The c1 accuracy figure compares the full-size GPU outputs with the oracle's outputs on a central
    crop, away from the crop border: there the crop's IUWT and linear convolution equal the full
    image's (checked here with the oracle on both sides, at a reduced size)."""
    import bench
    from oracle import convolution64 as conv
    from pymoresane_amd import synth
    n, S = 512, 4
    img, _ = synth.config_c1(n=n)
    psf = synth.make_gaussian_psf(2 * n, 6.0)
    det, sm = iuwt.iuwt_decomposition(img.astype(np.float64), S, store_smoothed=True)
    out = conv.fft_convolve(img, conv.psf_spectrum(psf), "linear")
    _, ref = bench.cpu_baseline(img, S, sample_n=256)
    err = bench.c1_accuracy(ref, det, sm, out, margin=64)  # S=4 reaches 14 + 32 pixels
    assert err["iuwt_detail_rel_l2_max"] < 1e-14
    assert err["iuwt_smoothed_rel_l2"] < 1e-14
    assert err["fft_convolve_rel_l2"] < 1e-13


def test_mp_restatement_matches_the_serial_oracle():
    """The strip-parallel 'mp' restatement (timed as the mp CPU baseline) computes the same transform as
    the serial float64 oracle up to its float32 shared array (iuwt.py:193-382 vs 66-191), including a
    strip count that does not divide the side."""
    from oracle import mp_iuwt
    x = np.random.default_rng(12).standard_normal((60, 60)).astype(np.float32)
    det, sm = mp_iuwt.mp_iuwt_decomposition(x, 4, 1, store_smoothed=True, core_count=3)
    rdet, rsm = iuwt.iuwt_decomposition(x, 4, 1, store_smoothed=True)
    assert np.max(np.abs(det - rdet)) < 1e-5 and np.max(np.abs(sm - rsm)) < 1e-5
    rec = mp_iuwt.mp_iuwt_recomposition(det.astype(np.float32), 1, core_count=3, smoothed_array=sm)
    assert np.max(np.abs(rec - iuwt.iuwt_recomposition(rdet, 1, rsm))) < 1e-5
