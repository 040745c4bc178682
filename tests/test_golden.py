"""Golden fixtures (tests/golden/): oracle regression on the CPU, FITS round trip, and the HIP path on the GPU."""

import json
import os
import sys

import numpy as np
import pytest

from oracle import convolution64 as oconv, iuwt64 as oiuwt, moresane64 as omain, toolbox64 as otools
from pymoresane_amd import fits

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF_TESTS = "/root/reference/tests"
TART_KW = dict(start_scale=1, stop_scale=20, sigma_level=4, loop_gain=0.2, tolerance=0.75, accuracy=1e-6,
               major_loop_miter=100, minor_loop_miter=50, conv_mode="circular", enforce_positivity=True)


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def trace_of(z):
    return [tuple(e) for e in json.loads(str(z["trace"]))]


def discrete(trace):
    out = []
    for e in trace:
        e = tuple(e)
        if e[0] == "scales":
            out.append(e[:5])
        elif e[0] == "minor":
            out.append(e[:4])
        elif e[0] == "update":
            out.append((e[0], e[1], bool(e[4])))
        else:
            out.append(e)
    return out


# ------------------------------------------------------------------------------------------ CPU
def test_oracle_reproduces_small_fixtures():
    g = load("oracle_small.npz")
    det, sm = oiuwt.iuwt_decomposition(g["img"], 4, 1, store_smoothed=True)
    np.testing.assert_allclose(det, g["det"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(sm, g["smoothed"], rtol=0, atol=1e-6)
    conv = oconv.fft_convolve(g["x"], oconv.psf_spectrum(g["psf"]), "linear")
    np.testing.assert_allclose(conv, g["conv"], rtol=0, atol=1e-5)
    np.testing.assert_array_equal(otools.estimate_threshold(g["planes"], 2, 3), g["thr"])
    ext, mask, labels = otools.source_extraction(g["planes"], 0.5, return_labels=True)
    np.testing.assert_array_equal(mask, g["mask"])
    np.testing.assert_array_equal(labels, g["labels"])


def test_tart_inputs_match_reference_fits():
    if not os.path.isdir(REF_TESTS):
        pytest.skip("reference tree not present")
    g = load("tart_inputs.npz")
    _, d = fits.read(os.path.join(REF_TESTS, "dirty_2017-10-15_20:33:35_UTC.fits"))
    _, b = fits.read(os.path.join(REF_TESTS, "beam_2017-10-15_20:33:35_UTC.fits"))
    np.testing.assert_array_equal(g["dirty"], d.astype(np.float32))
    np.testing.assert_array_equal(g["beam"], b.astype(np.float32))
    assert d.shape == (256, 256) and abs(b.max() - 0.00750732421875) < 1e-12


def test_oracle_reproduces_tart_moresane():
    g = load("tart_inputs.npz")
    z = load("tart_moresane.npz")
    ref = omain.OracleImage(g["dirty"], g["beam"])
    with np.errstate(divide="ignore"):
        ref.moresane_by_scale(**TART_KW)
    assert discrete(ref.trace) == discrete(trace_of(z))
    assert rel(ref.model, z["model"]) < 1e-6 and rel(ref.residual, z["residual"]) < 1e-6


def test_fits_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    a = rng.standard_normal((1, 1, 16, 16)).astype(np.float32)
    h = fits.Header([("CTYPE1", "RA---TAN", ""), ("CTYPE2", "DEC--TAN", ""), ("CDELT1", 0.5, ""),
                     ("CDELT2", 0.25, ""), ("OBJECT", "it's", "quote")])
    p = str(tmp_path / "x.fits")
    fits.write(p, a, h)
    h2, b = fits.read(p)
    np.testing.assert_array_equal(b, a)
    assert h2["CDELT2"] == 0.25 and h2["CTYPE1"] == "RA---TAN" and h2["OBJECT"] == "it's" and h2["NAXIS"] == 4
    assert os.path.getsize(p) % 2880 == 0


# ------------------------------------------------------------------------------------------ GPU
@pytest.mark.gpu
def test_hip_matches_small_fixtures(hip):
    from pymoresane_amd import iuwt, iuwt_convolution as conv, iuwt_toolbox as tools
    print("golden small fixtures"); sys.stdout.flush()
    g = load("oracle_small.npz")
    det, sm = iuwt.iuwt_decomposition(g["img"], 4, 1, store_smoothed=True)
    assert rel(det, g["det"]) < 1e-5 and rel(sm, g["smoothed"]) < 1e-5
    out = conv.fft_convolve(g["x"], conv.psf_spectrum(g["psf"]), conv_mode="linear")
    assert rel(out, g["conv"]) < 2e-6
    np.testing.assert_array_equal(tools.estimate_threshold(g["planes"], 2, 3), g["thr"])
    ext, mask, labels = tools.source_extraction(g["planes"], 0.5, return_labels=True)
    np.testing.assert_array_equal(mask, g["mask"])
    np.testing.assert_array_equal(labels, g["labels"])
    assert tools.snr_ratio(g["planes"], 0.9 * g["planes"]) == pytest.approx(float(g["snr"][0]), rel=1e-12)


@pytest.mark.gpu
def test_hip_fitsimage_tart_by_scale(hip, tmp_path):
    """The reference's own end-to-end case (tests/Makefile:8, runsane -ep) through FitsImage on the GPU."""
    from pymoresane_amd.main import FitsImage
    print("golden TART by-scale"); sys.stdout.flush()
    g = load("tart_inputs.npz")
    z = load("tart_moresane.npz")
    h = fits.Header([("CTYPE1", "RA---TAN", ""), ("CTYPE2", "DEC--TAN", ""), ("CDELT1", 0.703125, ""),
                     ("CDELT2", 0.703125, "")])
    dp, bp = str(tmp_path / "dirty.fits"), str(tmp_path / "beam.fits")
    fits.write(dp, g["dirty"], h)
    fits.write(bp, g["beam"], h)
    img = FitsImage(dp, bp)
    img.moresane_by_scale(**TART_KW)
    assert discrete(img.trace) == discrete(trace_of(z))
    assert rel(img.model, z["model"]) < 1e-4
    assert rel(img.residual, z["residual"]) < 1e-4
    # restore: the GPU convolution of the model with the fitted beam, plus the residual
    from pymoresane_amd.beam_fit import beam_fit
    params = img.restore()
    beam, params2 = beam_fit(img.psf_data, img.cdelt1, img.cdelt2)
    expect = oconv.fft_convolve(img.model, oconv.psf_spectrum(beam), "circular") + img.residual
    assert rel(img.restored, expect) < 1e-5
    assert params == params2 and img.img_hdr["BMAJ"] == pytest.approx(params[0])
    out = str(tmp_path / "restored.fits")
    img.save_fits(img.restored, out)
    _, back = fits.read(out)
    np.testing.assert_array_equal(back.reshape(256, 256), img.restored)
