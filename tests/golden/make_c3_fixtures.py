"""Float64 oracle runs of one job per small size class of the ragged batch, as a GPU-test fixture.

    python tests/golden/make_c3_fixtures.py          # writes tests/golden/c3_oracle.npz (~3 min, CPU)

Jobs 0 (256^2 dirty, 512^2 PSF) and 32 (512^2, 1024^2) of the 96-pair batch (seed 3000 + i, PSF FWHM
U(3, 8) px; moresane with S=5, gain 0.2, sigma 4, 100 major / 30 minor iterations, tolerance 0.75,
positivity, linear convolution).  Per job the fixture keeps the decision trace, the model (flat
indices and float64 values of its non-zero pixels) and the residual's norm and every 32nd row, which
cross-check the residual a test rebuilds from the model as dirty - H model.  The oracle runs with its
pinned fast backends (scipy correlate1d a trous, threaded scipy.fft).
"""

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

KW = dict(loop_gain=0.2, sigma_level=4, major_loop_miter=100, tolerance=0.75, enforce_positivity=True,
          conv_mode="linear", minor_loop_miter=30, accuracy=1e-6, scale_count=5)
JOBS = (0, 32)


def main():
    from oracle import convolution64, iuwt64, moresane64
    from pymoresane_amd import synth
    iuwt64.use_scipy_backend(True)
    convolution64.set_fft_workers(8)
    sides = synth.config_c3_sides()
    out = {"kw": np.array(json.dumps(KW)), "jobs": np.array(JOBS)}
    for i in JOBS:
        dirty, psf = synth.config_c3_pair(i, sides[i])
        img = moresane64.OracleImage(dirty, psf)
        t0 = time.time()
        img.moresane(**KW)
        print("job %d (%d^2): %.0f s, %d decisions" % (i, sides[i], time.time() - t0, len(img.trace)))
        idx = np.flatnonzero(img.model)
        out["j%d_model_idx" % i] = idx.astype(np.int32)
        out["j%d_model_val" % i] = img.model.ravel()[idx]
        out["j%d_trace" % i] = np.array(json.dumps([list(e) for e in img.trace]))
        out["j%d_residual_norm" % i] = np.float64(np.linalg.norm(img.residual))
        out["j%d_residual_rows" % i] = img.residual[::32].copy()
    path = os.path.join(HERE, "c3_oracle.npz")
    np.savez_compressed(path, **out)
    print("wrote %s (%d bytes)" % (path, os.path.getsize(path)))


if __name__ == "__main__":
    main()
