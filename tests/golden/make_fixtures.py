"""Write the golden fixtures in tests/golden/ (run in the build container, which holds the reference tree).

* tart_inputs.npz : the reference's own end-to-end test inputs (tests/dirty_2017-10-15_20:33:35_UTC.fits,
  tests/beam_2017-10-15_20:33:35_UTC.fits, used by tests/Makefile:8), read with pymoresane_amd.fits and
  cast to float32 as DataImage does (main.py:34-35).
* oracle_small.npz : float64-oracle outputs for small seeded inputs (IUWT, FFT convolution, MAD
  thresholds, extraction masks/labels, SNR) -- regression vectors for the oracle (CPU) and parity
  targets for the HIP path (GPU).
* tart_moresane.npz : the oracle's moresane_by_scale on the TART pair with the reference CLI defaults
  (runsane -ep: by-scale 1..20, sigma 4, gain 0.2, tolerance 0.75, 100 major / 50 minor iterations,
  circular convolution, positivity).
"""

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import convolution64 as oconv, iuwt64 as oiuwt, moresane64 as omain, toolbox64 as otools  # noqa: E402
from pymoresane_amd import fits  # noqa: E402

REF_TESTS = "/root/reference/tests"
TART_KW = dict(start_scale=1, stop_scale=20, sigma_level=4, loop_gain=0.2, tolerance=0.75, accuracy=1e-6,
               major_loop_miter=100, minor_loop_miter=50, conv_mode="circular", enforce_positivity=True)


def small_inputs():
    rng = np.random.default_rng(2024)
    img = rng.standard_normal((64, 64)).astype(np.float32)
    psf = rng.standard_normal((64, 64)).astype(np.float32)
    x = rng.standard_normal((32, 32)).astype(np.float32)
    planes = np.where(rng.random((3, 48, 48)) < 0.2, rng.standard_normal((3, 48, 48)), 0).astype(np.float32)
    return img, psf, x, planes


def main():
    _, dirty = fits.read(os.path.join(REF_TESTS, "dirty_2017-10-15_20:33:35_UTC.fits"))
    _, beam = fits.read(os.path.join(REF_TESTS, "beam_2017-10-15_20:33:35_UTC.fits"))
    dirty = dirty.astype(np.float32)
    beam = beam.astype(np.float32)
    np.savez_compressed(os.path.join(HERE, "tart_inputs.npz"), dirty=dirty, beam=beam)

    img, psf, x, planes = small_inputs()
    det, sm = oiuwt.iuwt_decomposition(img, 4, 1, store_smoothed=True)
    conv = oconv.fft_convolve(x, oconv.psf_spectrum(psf), "linear")
    thr = otools.estimate_threshold(planes, 2, 3)
    ext, mask, labels = otools.source_extraction(planes, 0.5, return_labels=True)
    snr = otools.snr_ratio(planes, 0.9 * planes)
    np.savez_compressed(os.path.join(HERE, "oracle_small.npz"), img=img, psf=psf, x=x, planes=planes,
                        det=det.astype(np.float32), smoothed=sm.astype(np.float32), conv=conv.astype(np.float32),
                        thr=thr, mask=mask.astype(np.uint8), labels=labels, snr=np.array([snr]))

    ref = omain.OracleImage(dirty, beam)
    ref.moresane_by_scale(**TART_KW)
    np.savez_compressed(os.path.join(HERE, "tart_moresane.npz"), model=ref.model.astype(np.float32),
                        residual=ref.residual.astype(np.float32),
                        trace=np.array(json.dumps([list(e) for e in ref.trace])))
    print("fixtures written:", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))


if __name__ == "__main__":
    main()
