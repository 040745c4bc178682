"""Float64 oracle run of the 2048^2 by-scale precision case, its committed fixture, and comparisons.

    python tests/golden/c2_precision.py ORACLE.npz            # run the oracle (~30 min, CPU)
    python tests/golden/c2_precision.py ORACLE.npz GPU.npz    # compare (oracle file must exist)
    python tests/golden/c2_precision.py ORACLE.npz --fixture tests/golden/c2_oracle.npz

The fixture keeps what a GPU test needs without the 32-minute oracle run: the decision trace, the
model (sparse: gaps between the flat indices of its non-zero pixels, and their values rounded to
float32), and, as a cross-check of the residual the test rebuilds from that model, the residual's
norm, sum and every 256th row.

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


def write_fixture(oracle_path, fixture_path):
    o = np.load(oracle_path)
    model, residual = o["model"], o["residual"]
    idx = np.flatnonzero(model)
    # flat indices as gaps (mostly 1: compresses to almost nothing), values rounded to float32 (6e-8
    # relative, far below the 1e-4 bar); the test rebuilds the model with a cumulative sum
    np.savez_compressed(fixture_path, shape=np.array(model.shape, dtype=np.int64), kw=np.array(json.dumps(KW)),
                        model_gaps=np.diff(idx, prepend=-1).astype(np.uint32),
                        model_val=model.ravel()[idx].astype(np.float32),
                        trace=o["trace"], residual_norm=np.float64(np.linalg.norm(residual)),
                        residual_sum=np.float64(residual.sum()), residual_rows=residual[::256].astype(np.float64))
    print("fixture: %d model pixels, %d bytes" % (idx.size, os.path.getsize(fixture_path)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("oracle")
    ap.add_argument("gpu", nargs="?")
    ap.add_argument("--perturb", type=float, default=0.0)
    ap.add_argument("--fixture", metavar="OUT_NPZ", help="write the committed GPU-test fixture from ORACLE")
    args = ap.parse_args()
    if not os.path.exists(args.oracle):
        run_oracle(args.oracle, args.perturb)
    if args.fixture:
        write_fixture(args.oracle, args.fixture)
    if args.gpu:
        compare(args.oracle, args.gpu)


if __name__ == "__main__":
    main()
