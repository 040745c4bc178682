"""Float64 oracle run of the 2048^2 by-scale precision case, and its comparison with a GPU result.

    python tests/golden/c2_precision.py ORACLE.npz            # run the oracle (~30 min, CPU)
    python tests/golden/c2_precision.py ORACLE.npz GPU.npz    # compare (oracle file must exist)

GPU.npz is written on an MI355X by ``python bench.py --precision GPU.npz`` (same generator,
seed 2: 2048^2 dirty, 4096^2 PSF, 2000 points + 50 blobs, noise 1e-4; moresane_by_scale 1..6
with loop_gain 0.2, sigma 4, tolerance 0.75, positivity, linear convolution).  The oracle uses
its pinned fast backends (scipy correlate1d a trous, threaded scipy.fft); with ``--perturb E``
the dirty image is multiplied by (1 + E * N(0,1)) first, to measure how far the float64
trajectory itself moves under input noise of that size.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

KW = dict(loop_gain=0.2, sigma_level=4, major_loop_miter=100, tolerance=0.75, enforce_positivity=True,
          conv_mode="linear", minor_loop_miter=30, accuracy=1e-6)


def run_oracle(out_path, perturb=0.0, workers=8):
    from oracle import convolution64, iuwt64, moresane64
    from pymoresane_amd import synth
    iuwt64.use_scipy_backend(True)
    convolution64.set_fft_workers(workers)
    dirty, psf = synth.config_c2(n=2048)
    if perturb:
        rng = np.random.default_rng(77)
        dirty = dirty.astype(np.float64) * (1.0 + perturb * rng.standard_normal(dirty.shape))
    img = moresane64.OracleImage(dirty, psf)
    t0 = time.time()
    img.moresane_by_scale(start_scale=1, stop_scale=6, **KW)
    print("oracle: %.0f s, %d decisions" % (time.time() - t0, len(img.trace)))
    trace = [[v.item() if hasattr(v, "item") else v for v in t] for t in img.trace]
    np.savez(out_path, model=img.model, residual=img.residual, trace=np.array(json.dumps(trace)))


def same_event(a, b, tol=1e-3):
    if a[0] != b[0] or len(a) != len(b):
        return False
    for x, y in zip(a[1:], b[1:]):
        if isinstance(x, float) or isinstance(y, float):
            if abs(x - y) > tol * max(1.0, abs(x)):
                return False
        elif x != y:
            return False
    return True


def compare(oracle_path, gpu_path):
    o, g = np.load(oracle_path), np.load(gpu_path)
    to, tg = json.loads(str(o["trace"])), json.loads(str(g["trace"]))
    first = next((i for i, (a, b) in enumerate(zip(to, tg)) if not same_event(a, b)), min(len(to), len(tg)))
    out = {"oracle_decisions": len(to), "gpu_decisions": len(tg), "first_divergence": first}
    for k in ("model", "residual"):
        a, b = o[k], g[k].astype(np.float64)
        out[k + "_rel_l2"] = float(np.linalg.norm(a - b) / np.linalg.norm(a))
        out[k + "_max_abs"] = float(np.abs(a - b).max())
    print(json.dumps(out))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("oracle")
    ap.add_argument("gpu", nargs="?")
    ap.add_argument("--perturb", type=float, default=0.0)
    args = ap.parse_args()
    if not os.path.exists(args.oracle):
        run_oracle(args.oracle, args.perturb)
    if args.gpu:
        compare(args.oracle, args.gpu)


if __name__ == "__main__":
    main()
