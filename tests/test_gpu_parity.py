"""HIP path vs the float64 oracle, through the C ABI (via the reference-shaped Python API).

Tolerances: float32 bulk arithmetic against the float64 oracle -- relative
L2 <= 1e-5 for single IUWT/FFT operations, <= 1e-4 for deconvolved models
and residuals.  Integer/index results (masks, labels, medians of identical
inputs, thresholded planes) are compared bit-exactly on identical inputs.
"""

import sys

import numpy as np
import pytest

from oracle import convolution64 as oconv
from oracle import iuwt64 as oiuwt
from oracle import moresane64 as omain
from oracle import toolbox64 as otools
from pymoresane_amd import iuwt, iuwt_convolution as conv, iuwt_toolbox as tools, synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg)
    sys.stdout.flush()


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    dt = np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(b) else np.float64
    a = a.astype(dt)
    b = b.astype(dt)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


# ------------------------------------------------------------------------------------------ IUWT
@pytest.mark.parametrize("shape,scales,adjust", [
    ((64, 64), 5, 0), ((64, 64), 5, 2), ((48, 80), 4, 1), ((40, 40), 4, 0), ((256, 256), 7, 0),
    ((256, 256), 4, 3), ((96, 32), 4, 0), ((512, 512), 6, 0),
    # planes above 256^2 pixels take the row-walk kernel
    ((1024, 1024), 5, 0), ((768, 1040), 4, 1), ((1040, 600), 6, 2),
])
def test_iuwt_decomposition_matches_oracle(hip, shape, scales, adjust):
    say("iuwt decomposition %s S=%d adj=%d" % (shape, scales, adjust))
    x = np.random.default_rng(1).standard_normal(shape).astype(np.float32)
    det, sm = iuwt.iuwt_decomposition(x, scales, adjust, store_smoothed=True)
    rdet, rsm = oiuwt.iuwt_decomposition(x, scales, adjust, store_smoothed=True)
    assert det.shape == rdet.shape and det.dtype == np.float32
    for k in range(det.shape[0]):
        assert rel(det[k], rdet[k]) < 1e-5, (k, rel(det[k], rdet[k]))
    assert rel(sm, rsm) < 1e-6


@pytest.mark.parametrize("shape,scales,adjust,smoothed", [
    ((64, 64), 5, 0, True), ((64, 64), 5, 2, False), ((48, 80), 4, 1, True), ((256, 256), 6, 0, False),
    ((32, 32), 1, 0, False), ((32, 32), 4, 3, False),
    ((768, 1040), 4, 1, True), ((1024, 1024), 5, 0, False), ((1040, 600), 3, 2, False),
])
def test_iuwt_recomposition_matches_oracle(hip, shape, scales, adjust, smoothed):
    say("iuwt recomposition %s S=%d adj=%d smoothed=%s" % (shape, scales, adjust, smoothed))
    rng = np.random.default_rng(2)
    det = rng.standard_normal((scales - adjust,) + shape).astype(np.float32)
    sm = rng.standard_normal(shape).astype(np.float32) if smoothed else None
    got = iuwt.iuwt_recomposition(det, adjust, smoothed_array=sm)
    ref = oiuwt.iuwt_recomposition(det, adjust, sm)
    assert rel(got, ref) < 1e-5


@pytest.mark.parametrize("shape,scales,adjust,smoothed", [
    ((1024, 1024), 6, 0, False), ((768, 1040), 5, 1, True), ((1040, 600), 5, 0, True),
    ((600, 1040), 4, 2, False), ((300, 2048), 6, 0, False), ((2048, 300), 7, 1, True),
])
def test_iuwt_recomposition_two_scale_launches_are_exact(hip, shape, scales, adjust, smoothed, monkeypatch):
    """Consecutive detail scales recomposed in one launch give the same bits as one launch per scale
    (the fused walk accumulates mirrored border taps in the order the one-scale kernel uses)."""
    say("iuwt recomposition pairs vs single scales %s S=%d adj=%d smoothed=%s" % (shape, scales, adjust, smoothed))
    rng = np.random.default_rng(5)
    det = rng.standard_normal((scales - adjust,) + shape).astype(np.float32)
    sm = rng.standard_normal(shape).astype(np.float32) if smoothed else None
    monkeypatch.setenv("PMS_IUWT_PAIR", "1")
    fused = iuwt.iuwt_recomposition(det, adjust, smoothed_array=sm)
    monkeypatch.setenv("PMS_IUWT_PAIR", "0")
    single = iuwt.iuwt_recomposition(det, adjust, smoothed_array=sm)
    np.testing.assert_array_equal(fused, single)
    assert rel(fused, oiuwt.iuwt_recomposition(det, adjust, sm)) < 1e-5


@pytest.mark.parametrize("shape,scales", [((1056, 784), 5), ((2048, 300), 6)])
def test_iuwt_autotuned_launches_are_exact(hip, shape, scales, monkeypatch, tmp_path):
    """The autotuner only changes how the walks are cut (workgroup targets, recomposition pairing):
    tuned and default launches give the same bits, and the choice lands in the JSON cache."""
    say("iuwt autotune %s S=%d" % (shape, scales))
    x = np.random.default_rng(6).standard_normal(shape).astype(np.float32)
    cache = tmp_path / "tune.json"
    monkeypatch.setenv("PMS_AUTOTUNE_CACHE", str(cache))
    monkeypatch.setenv("PMS_AUTOTUNE", "1")
    det_t, sm_t = iuwt.iuwt_decomposition(x, scales, store_smoothed=True)
    rec_t = iuwt.iuwt_recomposition(det_t, 0, smoothed_array=sm_t)
    monkeypatch.setenv("PMS_AUTOTUNE", "0")
    det_d, sm_d = iuwt.iuwt_decomposition(x, scales, store_smoothed=True)
    rec_d = iuwt.iuwt_recomposition(det_d, 0, smoothed_array=sm_d)
    np.testing.assert_array_equal(det_t, det_d)
    np.testing.assert_array_equal(sm_t, sm_d)
    np.testing.assert_array_equal(rec_t, rec_d)
    import json
    entries = json.loads(cache.read_text())
    tag = "%dx%d" % shape
    assert any(("decompose " + tag) in k for k in entries) and any(("recompose " + tag) in k for k in entries)
    for v in entries.values():
        assert v["target_blocks"] in (512, 1024, 2048, 4096) and v["pair"] in (0, 1, 2)


def test_moresane_two_scale_recomposition_is_exact(hip, monkeypatch):
    """Masked recompositions inside the minor loop: the whole deconvolution is bit-identical with and
    without two-scale launches."""
    say("moresane 1024^2, S=4, two-scale recomposition on/off")
    from pymoresane_amd.main import DataImage
    dirty, psf = synth.make_pair(1024, 31, n_points=30, n_blobs=3)
    kw = dict(scale_count=4, loop_gain=0.2, major_loop_miter=2, minor_loop_miter=12, enforce_positivity=True)
    out = []
    for flag in ("1", "0"):
        monkeypatch.setenv("PMS_IUWT_PAIR", flag)
        img = DataImage(dirty, psf, None, 1.0, 1.0)
        img.moresane(**kw)
        out.append(img)
    np.testing.assert_array_equal(out[0].model, out[1].model)
    np.testing.assert_array_equal(out[0].residual, out[1].residual)
    assert out[0].trace == out[1].trace


@pytest.mark.parametrize("shape,scales,adjust", [
    ((1024, 1024), 6, 0), ((300, 260), 5, 0), ((257, 333), 4, 1), ((768, 1040), 5, 1), ((160, 420), 5, 0),
    ((1040, 600), 6, 2), ((2048, 2048), 7, 0),
])
def test_iuwt_two_scale_decomposition_is_exact(hip, shape, scales, adjust, monkeypatch):
    """Pairs of detail scales decomposed in one launch (the intermediate smoothed plane only in LDS)
    give the same bits as one launch per scale, borders included."""
    say("iuwt decomposition pairs vs single scales %s S=%d adj=%d" % (shape, scales, adjust))
    x = np.random.default_rng(9).standard_normal(shape).astype(np.float32)
    monkeypatch.setenv("PMS_AUTOTUNE", "0")
    monkeypatch.setenv("PMS_IUWT_SMALL_LAUNCH", "0")  # row walks for every plane above 256^2
    monkeypatch.setenv("PMS_IUWT_DEC2", "1")
    det_f, sm_f = iuwt.iuwt_decomposition(x, scales, adjust, store_smoothed=True)
    monkeypatch.setenv("PMS_IUWT_DEC2", "0")
    det_s, sm_s = iuwt.iuwt_decomposition(x, scales, adjust, store_smoothed=True)
    np.testing.assert_array_equal(det_f, det_s)
    np.testing.assert_array_equal(sm_f, sm_s)
    rdet, rsm = oiuwt.iuwt_decomposition(x, scales, adjust, store_smoothed=True)
    for k in range(det_f.shape[0]):
        assert rel(det_f[k], rdet[k]) < 1e-5, (k, rel(det_f[k], rdet[k]))
    assert rel(sm_f, rsm) < 1e-6


def test_moresane_two_scale_decomposition_is_exact(hip, monkeypatch):
    """Masked decompositions inside the minor loop: the whole deconvolution is bit-identical with and
    without two-scale decomposition launches."""
    say("moresane 1024^2, S=4, two-scale decomposition on/off")
    from pymoresane_amd.main import DataImage
    dirty, psf = synth.make_pair(1024, 32, n_points=30, n_blobs=3)
    kw = dict(scale_count=4, loop_gain=0.2, major_loop_miter=2, minor_loop_miter=12, enforce_positivity=True)
    out = []
    for flag in ("1", "0"):
        monkeypatch.setenv("PMS_IUWT_DEC2", flag)
        img = DataImage(dirty, psf, None, 1.0, 1.0)
        img.moresane(**kw)
        out.append(img)
    np.testing.assert_array_equal(out[0].model, out[1].model)
    np.testing.assert_array_equal(out[0].residual, out[1].residual)
    assert out[0].trace == out[1].trace


@pytest.mark.parametrize("shape,scales,adjust", [
    ((256, 256), 4, 0), ((256, 256), 8, 0), ((256, 256), 9, 1), ((255, 201), 5, 0), ((64, 64), 5, 1),
    ((40, 40), 4, 0), ((128, 512), 6, 2), ((512, 128), 5, 0), ((17, 23), 3, 0), ((200, 31), 3, 0),
])
def test_iuwt_small_plane_scale_pairs_are_exact(hip, shape, scales, adjust, monkeypatch):
    """Planes of at most 256^2 pixels take scale pairs in one launch (the intermediate plane only in
    LDS, border positions at unfolded coordinates): decomposition and recomposition give the same
    bits as one launch per scale, up to dilations of a quarter of the side and next to scales beyond
    the side."""
    say("iuwt small-plane pairs vs single scales %s S=%d adj=%d" % (shape, scales, adjust))
    x = np.random.default_rng(11).standard_normal(shape).astype(np.float32)
    out = []
    for flag in ("1", "0"):
        monkeypatch.setenv("PMS_IUWT_SMALL2", flag)
        det, sm = iuwt.iuwt_decomposition(x, scales, adjust, store_smoothed=True)
        det_only = iuwt.iuwt_decomposition(x, scales, adjust, store_smoothed=False)
        rec = iuwt.iuwt_recomposition(det, adjust, smoothed_array=sm)
        rec_d = iuwt.iuwt_recomposition(det, adjust)
        out.append((det, sm, det_only, rec, rec_d))
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, b)
    rdet, rsm = oiuwt.iuwt_decomposition(x, scales, adjust, store_smoothed=True)
    for k in range(rdet.shape[0]):
        assert rel(out[0][0][k], rdet[k]) < 1e-5, (k, rel(out[0][0][k], rdet[k]))
    assert rel(out[0][1], rsm) < 1e-6
    assert rel(out[0][4], oiuwt.iuwt_recomposition(rdet, adjust, None)) < 1e-5


def test_moresane_small_plane_scale_pairs_are_exact(hip, monkeypatch):
    """Masked decompositions and recompositions of 256^2 planes inside the minor loop: the whole
    deconvolution is bit-identical with and without two-scale launches."""
    say("moresane 256^2, S=4, small-plane scale pairs on/off")
    from pymoresane_amd.main import DataImage
    dirty, psf = synth.make_pair(256, 33, n_points=30, n_blobs=3)
    kw = dict(scale_count=4, loop_gain=0.2, major_loop_miter=4, minor_loop_miter=30, enforce_positivity=True)
    out = []
    for flag in ("1", "0"):
        monkeypatch.setenv("PMS_IUWT_SMALL2", flag)
        img = DataImage(dirty, psf, None, 1.0, 1.0)
        img.moresane(**kw)
        out.append(img)
    np.testing.assert_array_equal(out[0].model, out[1].model)
    np.testing.assert_array_equal(out[0].residual, out[1].residual)
    assert out[0].trace == out[1].trace


def test_iuwt_mid_planes_same_on_both_kernel_families(hip, monkeypatch):
    """Planes of up to 512^2 pixels take the direct-stencil kernels when a launch holds few of them and
    the row walks otherwise (PMS_IUWT_SMALL_LAUNCH=0 forces the walks): both give the same bits, for
    decompositions, recompositions and a whole masked deconvolution."""
    from pymoresane_amd.main import DataImage
    dirty, psf = synth.make_pair(512, 41, n_points=30, n_blobs=3)
    out = []
    for flag in (None, "0"):
        if flag is None:
            monkeypatch.delenv("PMS_IUWT_SMALL_LAUNCH", raising=False)
        else:
            monkeypatch.setenv("PMS_IUWT_SMALL_LAUNCH", flag)
        res = []
        for shape, scales, adjust in [((512, 512), 6, 0), ((384, 512), 5, 1), ((300, 500), 5, 0)]:
            say("iuwt %s S=%d adj=%d, kernel family %s" % (shape, scales, adjust, flag or "by launch size"))
            x = np.random.default_rng(3).standard_normal(shape).astype(np.float32)
            det, sm = iuwt.iuwt_decomposition(x, scales, adjust, store_smoothed=True)
            res += [det, sm, iuwt.iuwt_recomposition(det, adjust, smoothed_array=sm), iuwt.iuwt_recomposition(det, adjust)]
        say("moresane 512^2 S=5, kernel family %s" % (flag or "by launch size"))
        img = DataImage(dirty, psf, None, 1.0, 1.0)
        img.moresane(scale_count=5, loop_gain=0.2, major_loop_miter=3, minor_loop_miter=30, enforce_positivity=True)
        res += [img.model, img.residual, repr(img.trace)]
        out.append(res)
    for a, b in zip(out[0], out[1]):
        if isinstance(a, str):
            assert a == b
        else:
            np.testing.assert_array_equal(a, b)


def test_iuwt_round_trip_at_full_size(hip):
    """4096^2, S=7, store_smoothed: decomposition then recomposition reproduces the input."""
    say("iuwt round trip 4096^2 S=7")
    img, _ = synth.config_c1(n=4096)
    det, sm = iuwt.iuwt_decomposition(img, 7, store_smoothed=True, store_on_gpu=True)
    rec = iuwt.iuwt_recomposition(det, 0, smoothed_array=sm)
    assert rel(rec, img) < 1e-6
    # linearity: W(2a) == 2 W(a) exactly in binary floating point
    d2 = iuwt.iuwt_decomposition(2 * img[:1024, :1024], 5)
    d1 = iuwt.iuwt_decomposition(img[:1024, :1024], 5)
    np.testing.assert_array_equal(d2, 2 * d1)


def test_iuwt_round_trip_at_the_largest_planes(hip):
    """16384^2 (2^28 pixels, 1 GiB per plane, just under the 2^29-pixel limit of the row walks' 32-bit
    byte offsets): decomposition then recomposition reproduces the input, and the finest detail plane
    matches the oracle on a corner tile computed from the corner alone (the 9-tap filter at scale 0
    reaches 4 pixels, so rows/columns below 60 of a 64-pixel tile see only tile pixels or the
    image's own reflected border)."""
    say("iuwt round trip 16384^2 S=4")
    n = 16384
    img = np.random.default_rng(14).standard_normal((n, n), dtype=np.float32)
    det, sm = iuwt.iuwt_decomposition(img, 4, store_smoothed=True, store_on_gpu=True)
    rec = iuwt.iuwt_recomposition(det, 0, smoothed_array=sm)
    assert rel(rec, img) < 1e-6
    w0 = det.get()[0, :64, :64]
    ref = oiuwt.iuwt_decomposition(img[:64, :64], 1)[0]
    assert rel(w0[:60, :60], ref[:60, :60]) < 1e-6


def test_iuwt_constant_image_and_device_inputs(hip):
    say("iuwt constant image")
    c = np.full((128, 128), 2.5, np.float32)
    det, sm = iuwt.iuwt_decomposition(c, 6, store_smoothed=True)
    assert np.max(np.abs(det)) < 1e-6
    assert np.max(np.abs(sm - 2.5)) < 1e-6


@pytest.mark.parametrize("shape,scales,adjust", [((16, 16), 6, 0), ((48, 80), 7, 1), ((512, 384), 9, 0),
                                                  ((1024, 300), 10, 2)])
def test_iuwt_scales_beyond_the_side_match_oracle(hip, shape, scales, adjust):
    """Scales whose taps exceed a side (2^(s+1) > side): the reference's clipped slices read that axis
    reversed (iuwt.py:166-187); decomposition and recomposition match the oracle, which is pinned to
    the reference's slice arithmetic, and the round trip is exact."""
    say("iuwt over-long scales %s S=%d adj=%d" % (shape, scales, adjust))
    x = np.random.default_rng(2).standard_normal(shape).astype(np.float32)
    det, sm = iuwt.iuwt_decomposition(x, scales, adjust, store_smoothed=True)
    rdet, rsm = oiuwt.iuwt_decomposition(x, scales, adjust, store_smoothed=True)
    for k in range(det.shape[0]):
        assert rel(det[k], rdet[k]) < 1e-5, (k, rel(det[k], rdet[k]))
    assert rel(sm, rsm) < 1e-6
    rec = iuwt.iuwt_recomposition(det, adjust)
    assert rel(rec, oiuwt.iuwt_recomposition(rdet, adjust)) < 1e-5
    if adjust == 0:
        assert rel(iuwt.iuwt_recomposition(det, 0, smoothed_array=sm), x) < 1e-5


# ------------------------------------------------------------------------------------------- FFT
@pytest.mark.parametrize("n,mode", [(4, "linear"), (8, "linear"), (16, "circular"), (64, "linear"),
                                    (256, "linear"), (256, "circular"), (512, "linear"), (1024, "circular")])
def test_fft_convolve_matches_oracle(hip, n, mode):
    say("fft_convolve n=%d %s" % (n, mode))
    rng = np.random.default_rng(3)
    x = rng.standard_normal((n, n)).astype(np.float32)
    p = 2 * n if mode == "linear" else n
    psf = rng.standard_normal((p, p)).astype(np.float32)
    got = conv.fft_convolve(x, conv.psf_spectrum(psf), conv_mode=mode)
    ref = oconv.fft_convolve(x, oconv.psf_spectrum(psf), mode)
    assert got.dtype == np.float32 and got.shape == (n, n)
    assert rel(got, ref) < 2e-6, rel(got, ref)


# Frames whose side is not a power of two (the reference's numpy rfft2 takes any size,
# iuwt_convolution.py:64-73): mixed-radix transforms with every supported prime (2, 3, 5, 7, 11, 13).
@pytest.mark.parametrize("n,mode", [(6, "linear"), (48, "linear"), (52, "circular"), (88, "linear"),
                                    (96, "linear"), (96, "circular"), (98, "linear"), (160, "linear"),
                                    (160, "circular"), (480, "linear"), (500, "linear"), (1000, "linear"),
                                    (1000, "circular"), (3000, "linear"), (3000, "circular")])
def test_fft_convolve_any_even_side_matches_oracle(hip, n, mode):
    say("fft_convolve n=%d %s (mixed radix)" % (n, mode))
    rng = np.random.default_rng(31)
    x = rng.standard_normal((n, n)).astype(np.float32)
    p = 2 * n if mode == "linear" else n
    psf = rng.standard_normal((p, p)).astype(np.float32)
    got = conv.fft_convolve(x, conv.psf_spectrum(psf), conv_mode=mode)
    ref = oconv.fft_convolve(x, oconv.psf_spectrum(psf), mode)
    assert got.dtype == np.float32 and got.shape == (n, n)
    assert rel(got, ref) < 5e-6, rel(got, ref)


@pytest.mark.parametrize("p", [12, 96, 200, 1000])
def test_psf_spectrum_any_even_side_round_trip(hip, p):
    """The device spectrum of a non-power-of-two frame, read back in numpy rfft2 layout, equals
    numpy's rfft2 of the frame, and a numpy spectrum converted to the library layout convolves like
    the device-built one."""
    say("psf spectrum p=%d (mixed radix)" % p)
    rng = np.random.default_rng(32)
    psf = rng.standard_normal((p, p)).astype(np.float32)
    s = conv.gpu_r2c_fft(psf)
    assert rel(s, np.fft.rfft2(psf.astype(np.float64))) < 2e-6
    n = p // 2
    x = rng.standard_normal((n, n)).astype(np.float32)
    a = conv.fft_convolve(x, conv.psf_spectrum(psf), conv_mode="linear")
    b = conv.fft_convolve(x, np.fft.rfft2(psf), conv_mode="linear")
    assert rel(a, b) < 2e-6


def test_fft_convolve_rejects_large_prime_factors(hip):
    x = np.zeros((34, 34), np.float32)  # P = 68 = 4 * 17
    with pytest.raises(Exception, match="prime factor"):
        conv.fft_convolve(x, conv.psf_spectrum(np.zeros((68, 68), np.float32)), conv_mode="linear")


def test_fft_convolve_accepts_numpy_spectrum_and_padded_psf(hip):
    say("fft_convolve numpy spectrum / padded psf")
    rng = np.random.default_rng(4)
    n = 128
    x = rng.standard_normal((n, n)).astype(np.float32)
    psf = synth.make_psf(n)  # same-size PSF: the linear path pads it (main.py:203-206)
    ref = oconv.fft_convolve(x, oconv.psf_spectrum(oconv.pad_array(psf)), "linear")
    got_dev = conv.fft_convolve(x, conv.psf_spectrum(psf, 2 * n), conv_mode="linear")
    got_np = conv.fft_convolve(x, np.fft.rfft2(conv.pad_array(psf)), conv_mode="linear")
    assert rel(got_dev, ref) < 2e-6 and rel(got_np, ref) < 2e-6
    # gpu_r2c_fft round trip to the numpy layout
    s = conv.gpu_r2c_fft(oconv.pad_array(psf).astype(np.float32))
    assert rel(s, np.fft.rfft2(oconv.pad_array(psf))) < 2e-6


def test_fft_convolve_full_size_linearity(hip):
    """4096^2 image with the 8192^2 Gaussian PSF: conv(a+b) == conv(a)+conv(b), and a delta returns the PSF."""
    say("fft_convolve 4096 linear, P=8192")
    img, psf = synth.config_c1(n=4096)
    spec = conv.psf_spectrum(psf)
    n = 4096
    delta = np.zeros((n, n), np.float32)
    delta[1000, 3000] = 1.0
    out = conv.fft_convolve(delta, spec, conv_mode="linear")
    expect = psf[n - 1000:2 * n - 1000, n - 3000:2 * n - 3000]  # out[i] = psf[i - u + N]
    assert np.max(np.abs(out - expect)) < 2e-6
    a = img
    b = np.roll(img, 17, axis=1)
    ca = conv.fft_convolve(a, spec)
    cb = conv.fft_convolve(b, spec)
    cab = conv.fft_convolve(a + b, spec)
    assert rel(cab, ca + cb) < 1e-6


def test_fft_convolve_largest_linear_frame(hip):
    """8192^2 image, linear mode: the 16384^2 frame, the largest the library takes.  A delta at (u, v)
    returns the PSF window psf[i - u + N, j - v + N] (the relation the 4096^2 case checks), and the
    convolution is linear."""
    say("fft_convolve 8192 linear, P=16384")
    n = 8192
    g = np.exp(-0.5 * ((np.arange(2 * n, dtype=np.float32) - n) / 2.5) ** 2).astype(np.float32)
    psf = np.outer(g, g)
    spec = conv.psf_spectrum(psf)
    delta = np.zeros((n, n), np.float32)
    u, v = 3001, 5000
    delta[u, v] = 1.0
    out = conv.fft_convolve(delta, spec, conv_mode="linear")
    assert np.max(np.abs(out - psf[n - u:2 * n - u, n - v:2 * n - v])) < 2e-6
    rng = np.random.default_rng(13)
    a = rng.standard_normal((n, n), dtype=np.float32)
    b = np.roll(a, 17, axis=0)
    assert rel(conv.fft_convolve(a + b, spec), conv.fft_convolve(a, spec) + conv.fft_convolve(b, spec)) < 1e-6


@pytest.mark.parametrize("n,mode", [(4096, "linear"), (8192, "circular")])
def test_fft_convolve_large_frame_matches_oracle(hip, n, mode):
    """Random image and PSF at the 8192^2 frame, linear (the 4096^2 microbenchmark) and circular: the
    frames whose row passes take two row groups per workgroup, against the float64 oracle."""
    say("fft_convolve n=%d %s vs oracle" % (n, mode))
    rng = np.random.default_rng(14)
    x = rng.standard_normal((n, n), dtype=np.float32)
    p = 2 * n if mode == "linear" else n
    psf = rng.standard_normal((p, p), dtype=np.float32)
    got = conv.fft_convolve(x, conv.psf_spectrum(psf), conv_mode=mode)
    ref = oconv.fft_convolve(x, oconv.psf_spectrum(psf), mode)
    assert rel(got, ref) < 2e-6, rel(got, ref)


@pytest.mark.parametrize("n", [64, 8192])
def test_fft_convolve_circular_delta_response(hip, n):
    """Circular mode (P = n) answers a delta at (a, b) with the PSF rolled by (a + n/2, b + n/2) (the
    circular convolution followed by the reference's fftshift); checked against the oracle at 64 and
    used at 8192, the largest circular frame."""
    say("fft_convolve circular delta n=%d" % n)
    rng = np.random.default_rng(9)
    psf = rng.random((n, n), dtype=np.float32)
    x = np.zeros((n, n), np.float32)
    a, b = n // 3, n // 5 + 1
    x[a, b] = 1.0
    expect = np.roll(psf, (a + n // 2, b + n // 2), axis=(0, 1))
    if n <= 256:
        ref = oconv.fft_convolve(x, oconv.psf_spectrum(psf), "circular")
        assert np.max(np.abs(ref - expect)) < 1e-9
    got = conv.fft_convolve(x, conv.psf_spectrum(psf), conv_mode="circular")
    assert np.max(np.abs(got - expect)) < 1e-5  # PSF values in [0, 1): a wrong roll or crop errs by O(1)


# ------------------------------------------------------------------------------------ threshold
@pytest.mark.parametrize("shape,ee,ie", [((64, 64), 0, 0), ((63, 64), 0, 0), ((64, 64), 5, 0), ((64, 64), 0, 7),
                                         ((64, 64), 3, 9), ((256, 256), 0, 0), ((1024, 1024), 0, 0),
                                         # non-square: the excluded square is centred on the row count
                                         ((48, 80), 0, 7), ((80, 48), 3, 9),
                                         # h*w % 4 != 0: planes after the first start unaligned
                                         ((33, 47), 0, 0), ((33, 47), 2, 5)])
def test_estimate_threshold_is_exact(hip, shape, ee, ie):
    say("estimate_threshold %s edge=%d int=%d" % (shape, ee, ie))
    rng = np.random.default_rng(5)
    w = rng.standard_normal((3,) + shape).astype(np.float32)
    w[1, :5, :5] = 0.0          # ties at zero
    w[2] = np.round(w[2] * 4) / 4  # many ties
    got = tools.estimate_threshold(w, ee, ie)
    ref = otools.estimate_threshold(w, ee, ie)
    np.testing.assert_array_equal(got, ref)


def test_apply_threshold_is_exact(hip):
    say("apply_threshold")
    rng = np.random.default_rng(6)
    w = rng.standard_normal((4, 96, 96)).astype(np.float32)
    thr = otools.estimate_threshold(w)
    w[0, 0, 0] = np.float32(4 * thr[0])  # on the threshold: dropped (strict >)
    got = tools.apply_threshold(w, thr, 4)
    ref = otools.apply_threshold(w, thr, 4)
    np.testing.assert_array_equal(got, ref.astype(np.float32))
    two_d = tools.apply_threshold(w[1], thr[1], 3)
    np.testing.assert_array_equal(two_d, otools.apply_threshold(w[1], thr[1], 3).astype(np.float32))


# ----------------------------------------------------------------------------------- extraction
def sparse_planes(seed, s, h, w, density=0.08):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((s, h, w)).astype(np.float32)
    keep = rng.random((s, h, w)) < density
    # grow blobs so objects span several pixels and scales overlap
    from scipy import ndimage
    keep = np.stack([ndimage.binary_dilation(k, iterations=1) for k in keep])
    return np.where(keep, v, 0.0).astype(np.float32)


@pytest.mark.parametrize("seed,shape,neg", [(0, (3, 64, 64), False), (1, (4, 128, 96), False),
                                            (2, (2, 64, 64), True), (3, (1, 33, 47), False),
                                            (4, (5, 256, 256), False), (5, (2, 100, 70), True),
                                            (6, (3, 600, 530), False)])
def test_source_extraction_bit_exact(hip, seed, shape, neg):
    say("source_extraction seed=%d %s neg=%s" % (seed, shape, neg))
    w = sparse_planes(seed, *shape)
    ext, mask, labels = tools.source_extraction(w, 0.5, neg_comp=neg, return_labels=True)
    rext, rmask, rlabels = otools.source_extraction(w, 0.5, neg_comp=neg, return_labels=True)
    np.testing.assert_array_equal(labels, rlabels)
    np.testing.assert_array_equal(mask, rmask)
    np.testing.assert_array_equal(ext, rext.astype(np.float32))


def test_source_extraction_components_spanning_tiles(hip):
    """Dense planes: components percolate across many 32x32 labelling tiles."""
    say("source_extraction dense planes")
    for seed, dens in ((7, 0.3), (8, 0.45)):
        w = sparse_planes(seed, 2, 300, 260, density=dens)
        ext, mask, labels = tools.source_extraction(w, 0.5, return_labels=True)
        rext, rmask, rlabels = otools.source_extraction(w, 0.5, return_labels=True)
        np.testing.assert_array_equal(labels, rlabels)
        np.testing.assert_array_equal(mask, rmask)
        np.testing.assert_array_equal(ext, rext.astype(np.float32))


def test_source_extraction_empty_and_full(hip):
    say("source_extraction empty/full")
    z = np.zeros((2, 32, 32), np.float32)
    ext, mask = tools.source_extraction(z, 0.75)
    assert not mask.any() and not ext.any()
    f = np.ones((1, 32, 32), np.float32)
    ext, mask = tools.source_extraction(f, 0.75)
    assert mask.all()


def test_snr_ratio(hip):
    say("snr_ratio")
    rng = np.random.default_rng(8)
    a = rng.standard_normal((3, 64, 64)).astype(np.float32)
    b = a + 0.01 * rng.standard_normal(a.shape).astype(np.float32)
    assert tools.snr_ratio(a, b) == pytest.approx(otools.snr_ratio(a, b), rel=1e-12)


# ------------------------------------------------------------------------------------ MORESANE
def discrete(trace):
    """Decision trace without the floating-point payloads."""
    out = []
    for e in trace:
        if e[0] == "scales":
            out.append(e[:5])
        elif e[0] == "minor":
            out.append(e[:4])
        elif e[0] == "update":
            out.append((e[0], e[1], e[4]))
        else:
            out.append(e)
    return out


def run_pair(dirty, psf, **kw):
    from pymoresane_amd.main import DataImage
    img = DataImage(dirty, psf, None, 1.0, 1.0)
    img.moresane(**kw)
    ref = omain.OracleImage(dirty, psf)
    ref.moresane(**kw)
    return img, ref


@pytest.mark.parametrize("case", ["c0", "mid", "walk"])
def test_moresane_incremental_snr_matches_direct(hip, case, monkeypatch):
    """The minor loop's SNR sums by linearity (M.W(H x) carried across iterations, updated with alpha
    times this iteration's M.W(H p)) against the direct convolution + decomposition of xn that
    main.py:413-424 writes: the same decisions, hence the same model and residual bits, and SNR
    values that agree to float32 rounding."""
    from pymoresane_amd.main import DataImage
    if case == "c0":
        dirty, psf = synth.config_c0()
        kw = dict(scale_count=4, loop_gain=0.2, sigma_level=4, major_loop_miter=30, tolerance=0.75,
                  enforce_positivity=True, conv_mode="linear", minor_loop_miter=30, accuracy=1e-6)
    else:
        # "mid": 512^2 planes, whose reference-plane decompositions take the row walks while the
        # others take the direct-stencil kernels
        dirty, psf = synth.make_pair(1024 if case == "walk" else 512, 33, n_points=30, n_blobs=3)
        kw = dict(scale_count=4, loop_gain=0.2, major_loop_miter=3, minor_loop_miter=12, enforce_positivity=True)
    say("moresane %s: incremental vs direct SNR" % case)
    out = []
    for flag in ("0", "1"):
        monkeypatch.setenv("PMS_SNR_DIRECT", flag)
        img = DataImage(dirty, psf, None, 1.0, 1.0)
        img.moresane(**kw)
        out.append(img)
    inc, direct = out
    assert discrete(inc.trace) == discrete(direct.trace)
    np.testing.assert_array_equal(inc.model, direct.model)
    np.testing.assert_array_equal(inc.residual, direct.residual)
    snr_i = [e[4] for e in inc.trace if e[0] == "minor"]
    snr_d = [e[4] for e in direct.trace if e[0] == "minor"]
    assert len(snr_i) > 0
    assert np.allclose(snr_i, snr_d, rtol=1e-5, atol=1e-6), (snr_i, snr_d)


def test_moresane_small_linear(hip):
    say("moresane 64^2 linear")
    dirty, psf = synth.make_pair(64, 21, n_points=8, n_blobs=1)
    img, ref = run_pair(dirty, psf, scale_count=3, loop_gain=0.2, major_loop_miter=20, enforce_positivity=True)
    assert discrete(img.trace) == discrete(ref.trace)
    assert rel(img.model, ref.model) < 1e-4
    assert rel(img.residual, ref.residual) < 1e-4


@pytest.mark.parametrize("case", ["noise_only", "all_zero"])
def test_moresane_without_sources(hip, case):
    """Inputs with nothing to deconvolve: noise alone (no significant coefficient at the first major
    iteration, then no work: the oracle's trace is nocoeff, nowork) and an all-zero image.  The trace,
    model and residual are the oracle's."""
    say("moresane 64^2 without sources: %s" % case)
    _, psf = synth.make_pair(64, 23, n_points=10, n_blobs=2)
    if case == "noise_only":
        dirty = np.random.default_rng(3).normal(0, 1e-3, (64, 64)).astype(np.float32)
    else:
        dirty = np.zeros((64, 64), np.float32)
    img, ref = run_pair(dirty, psf, scale_count=4, enforce_positivity=True)
    assert discrete(img.trace) == discrete(ref.trace)
    if case == "noise_only":
        assert [e[0] for e in ref.trace] == ["nocoeff", "nowork"]
    np.testing.assert_array_equal(img.model, ref.model.astype(np.float32))
    np.testing.assert_array_equal(img.residual, ref.residual.astype(np.float32))


def test_moresane_small_circular_options(hip):
    say("moresane 64^2 circular + options")
    dirty, psf = synth.make_pair(64, 22, psf_factor=1, n_points=8, n_blobs=1)
    kw = dict(scale_count=4, loop_gain=0.3, major_loop_miter=15, conv_mode="circular", edge_suppression=True,
              edge_excl=3, int_excl=4, neg_comp=True)
    img, ref = run_pair(dirty, psf, **kw)
    assert discrete(img.trace) == discrete(ref.trace)
    assert rel(img.model, ref.model) < 1e-4
    assert rel(img.residual, ref.residual) < 1e-4


@pytest.mark.parametrize("case", ["flux_edge_auto", "subregion_negcomp", "by_scale_from_2"])
def test_moresane_more_options(hip, case):
    """Options the other cases leave at their defaults: flux_threshold (shallow deconvolution,
    main.py:253-255), edge_suppression with an edge_offset, scale_count=None (log2 of the side,
    main.py:108-109), a subregion without a mask, neg_comp with a lower sigma level and tolerance, and a
    by-scale run that starts above scale 1: identical trace, model and residual within 1e-4."""
    say("moresane options: %s" % case)
    from pymoresane_amd.main import DataImage
    by_scale = False
    if case == "flux_edge_auto":
        dirty, psf = synth.make_pair(64, 26, n_points=10, n_blobs=1)
        kw = dict(scale_count=None, loop_gain=0.2, major_loop_miter=15, flux_threshold=2.5,
                  edge_suppression=True, edge_offset=2, enforce_positivity=True)
    elif case == "subregion_negcomp":
        dirty, psf = synth.make_pair(128, 27, n_points=12, n_blobs=2)
        kw = dict(subregion=64, scale_count=3, sigma_level=3, tolerance=0.6, loop_gain=0.2, major_loop_miter=10,
                  neg_comp=True)
    else:
        dirty, psf = synth.make_pair(64, 28, n_points=10, n_blobs=2)
        kw = dict(start_scale=2, stop_scale=4, loop_gain=0.2, major_loop_miter=10, enforce_positivity=True)
        by_scale = True
    img = DataImage(dirty, psf, None, 1.0, 1.0)
    ref = omain.OracleImage(dirty, psf)
    if by_scale:
        img.moresane_by_scale(**kw)
        ref.moresane_by_scale(**kw)
    else:
        img.moresane(**kw)
        ref.moresane(**kw)
    assert len(ref.trace) > 0
    if case == "flux_edge_auto":  # the flux threshold ends the major loop before its iteration cap
        assert sum(1 for e in ref.trace if e[0] == "update") < kw["major_loop_miter"]
    assert discrete(img.trace) == discrete(ref.trace)
    assert rel(img.model, ref.model) < 1e-4
    assert rel(img.residual, ref.residual) < 1e-4


@pytest.mark.parametrize("sub", [None, 64])
def test_moresane_with_mask(hip, sub):
    """A user mask (main.py:283-285: the decomposition of dirty*mask[subregion] replaces the thresholded
    one), with and without a subregion: identical trace, model and residual within 1e-4."""
    say("moresane 128^2 with a mask, subregion %s" % sub)
    from pymoresane_amd.main import DataImage
    dirty, psf = synth.make_pair(128, 25, n_points=12, n_blobs=2)
    yy, xx = np.mgrid[:128, :128]
    mask = (((yy - 60) ** 2 + (xx - 70) ** 2) < 40 ** 2).astype(np.float32)
    kw = dict(scale_count=4, loop_gain=0.2, major_loop_miter=12, enforce_positivity=True)
    if sub:
        kw["subregion"] = sub
    img = DataImage(dirty, psf, mask, 1.0, 1.0)
    img.moresane(**kw)
    ref = omain.OracleImage(dirty, psf, mask)
    ref.moresane(**kw)
    assert discrete(img.trace) == discrete(ref.trace)
    assert rel(img.model, ref.model) < 1e-4
    assert rel(img.residual, ref.residual) < 1e-4


@pytest.mark.parametrize("n,mode", [(96, "linear"), (160, "circular"), (1000, "linear")])
def test_moresane_any_even_side(hip, n, mode):
    """Deconvolution of images whose side is not a power of two (FFT frames 192, 160 and 2000):
    identical decision trace, model and residual within 1e-4 of the float64 oracle."""
    say("moresane %d^2 %s (mixed-radix frames)" % (n, mode))
    dirty, psf = synth.make_pair(n, 40 + n, psf_factor=2 if mode == "linear" else 1, n_points=12, n_blobs=2)
    kw = dict(scale_count=4, loop_gain=0.2, major_loop_miter=8 if n < 1000 else 3, minor_loop_miter=20,
              enforce_positivity=True, conv_mode=mode)
    img, ref = run_pair(dirty, psf, **kw)
    assert any(e[0] == "update" for e in ref.trace)
    assert discrete(img.trace) == discrete(ref.trace)
    assert rel(img.model, ref.model) < 1e-4
    assert rel(img.residual, ref.residual) < 1e-4


@pytest.mark.parametrize("by_scale", [False, True])
def test_moresane_small_subregion_automatic_scales(hip, by_scale):
    """A 64^2 subregion of a 256^2 image with scale_count=None: the scale count comes from the whole
    image (log2(256) - 1 = 7, main.py:108-109), so the coarsest scales of the subregion have taps
    longer than its side; by scale, the run goes on through those scale counts.  Identical trace,
    model and residual within 1e-4 of the oracle."""
    say("moresane 256^2, subregion 64, automatic scales (by_scale=%s)" % by_scale)
    from pymoresane_amd.main import DataImage
    dirty, psf = synth.make_pair(256, 61, n_points=20, n_blobs=2)
    kw = dict(subregion=64, loop_gain=0.2, major_loop_miter=6, minor_loop_miter=20, enforce_positivity=True)
    img = DataImage(dirty, psf, None, 1.0, 1.0)
    ref = omain.OracleImage(dirty, psf)
    if by_scale:
        img.moresane_by_scale(start_scale=5, stop_scale=8, **kw)
        ref.moresane_by_scale(start_scale=5, stop_scale=8, **kw)
    else:
        img.moresane(scale_count=None, **kw)
        ref.moresane(scale_count=None, **kw)
    assert len(ref.trace) > 0
    assert discrete(img.trace) == discrete(ref.trace)
    assert rel(img.model, ref.model) < 1e-4
    assert rel(img.residual, ref.residual) < 1e-4


def test_moresane_config_c0(hip):
    """256^2 dirty, 512^2 PSF, S=4, gain 0.2, sigma 4, 100 major, tol 0.75, positivity, linear."""
    say("moresane c0")
    dirty, psf = synth.config_c0()
    kw = dict(scale_count=4, loop_gain=0.2, sigma_level=4, major_loop_miter=100, tolerance=0.75,
              enforce_positivity=True, conv_mode="linear", minor_loop_miter=30, accuracy=1e-6)
    img, ref = run_pair(dirty, psf, **kw)
    say("c0 model rel %.3e residual rel %.3e" % (rel(img.model, ref.model), rel(img.residual, ref.residual)))
    assert discrete(img.trace) == discrete(ref.trace)
    assert rel(img.model, ref.model) < 1e-4
    assert rel(img.residual, ref.residual) < 1e-4


def test_moresane_row_walk_size(hip):
    """1024^2: the IUWT planes take the row-walk kernel (masked/reference variants in the CG loop)."""
    say("moresane 1024^2, S=3, 2 major iterations")
    dirty, psf = synth.make_pair(1024, 24, n_points=30, n_blobs=3)
    kw = dict(scale_count=3, loop_gain=0.2, major_loop_miter=2, minor_loop_miter=12, enforce_positivity=True)
    img, ref = run_pair(dirty, psf, **kw)
    say("1024 model rel %.3e residual rel %.3e" % (rel(img.model, ref.model), rel(img.residual, ref.residual)))
    assert discrete(img.trace) == discrete(ref.trace)
    assert rel(img.model, ref.model) < 1e-4
    assert rel(img.residual, ref.residual) < 1e-4


def test_moresane_by_scale_small(hip):
    say("moresane_by_scale 64^2")
    from pymoresane_amd.main import DataImage
    dirty, psf = synth.make_pair(64, 23, n_points=10, n_blobs=2)
    kw = dict(loop_gain=0.2, major_loop_miter=10, enforce_positivity=True)
    img = DataImage(dirty, psf, None, 1.0, 1.0)
    img.moresane_by_scale(start_scale=1, stop_scale=4, **kw)
    ref = omain.OracleImage(dirty, psf)
    ref.moresane_by_scale(start_scale=1, stop_scale=4, **kw)
    assert discrete(img.trace) == discrete(ref.trace)
    assert rel(img.model, ref.model) < 1e-4
    assert rel(img.residual, ref.residual) < 1e-4


def test_graph_replay_equals_eager(hip, monkeypatch):
    """The minor-loop iterations replayed from captured HIP graphs (with eager work and host syncs between
    replays) give bit-identical models, residuals and traces to launching every kernel eagerly."""
    say("graph replay vs eager, c0")
    from pymoresane_amd.main import DataImage
    dirty, psf = synth.config_c0()
    kw = dict(scale_count=4, loop_gain=0.2, sigma_level=4, major_loop_miter=30, tolerance=0.75,
              enforce_positivity=True, conv_mode="linear", minor_loop_miter=30, accuracy=1e-6)
    monkeypatch.setenv("PMS_CG_GRAPH", "0")
    eager = DataImage(dirty, psf, None, 1.0, 1.0)
    eager.moresane(**kw)
    monkeypatch.setenv("PMS_CG_GRAPH", "1")
    graph = DataImage(dirty, psf, None, 1.0, 1.0)
    graph.moresane(**kw)
    np.testing.assert_array_equal(graph.model, eager.model)
    np.testing.assert_array_equal(graph.residual, eager.residual)
    assert graph.trace == eager.trace


@pytest.mark.parametrize("graph", ["0", "1"])
def test_iterations_per_step_change_no_bit(hip, monkeypatch, graph):
    """Minor-loop iterations queued K per host step, their exit rules applied on the device (a job's
    iterations after its exit do nothing, the host replays the iterate swaps): K = 1, 2, 3, 8 give
    the same models, residuals and traces bit for bit, eagerly and from captured graphs, plain and
    by scale (whose minor loops end by every rule)."""
    say("iterations per step, graph=%s" % graph)
    from pymoresane_amd.main import DataImage
    dirty, psf = synth.config_c0()
    kw = dict(scale_count=4, loop_gain=0.2, sigma_level=4, major_loop_miter=30, tolerance=0.75,
              enforce_positivity=True, conv_mode="linear", minor_loop_miter=30, accuracy=1e-6)
    d2, p2 = synth.make_pair(128, 71, n_points=14, n_blobs=2)
    monkeypatch.setenv("PMS_CG_GRAPH", graph)
    out = []
    for k in ("1", "2", "3", "8"):
        monkeypatch.setenv("PMS_CG_ITERS", k)
        a = DataImage(dirty, psf, None, 1.0, 1.0)
        a.moresane(**kw)
        b = DataImage(d2, p2, None, 1.0, 1.0)
        b.moresane_by_scale(start_scale=1, stop_scale=5, loop_gain=0.2, major_loop_miter=10, minor_loop_miter=7,
                            enforce_positivity=True)
        out.append((a, b))
    codes = {e[3] for _, b in out for e in b.trace if e[0] == "minor"}
    assert len(codes) >= 3, codes  # several exit rules, the iteration cap among them
    for a, b in out[1:]:
        for x, y in ((a, out[0][0]), (b, out[0][1])):
            np.testing.assert_array_equal(x.model, y.model)
            np.testing.assert_array_equal(x.residual, y.residual)
            assert x.trace == y.trace
