"""The batched and served workloads at full size.

* Ragged batch: all 96 dirty/PSF pairs (32 x 256^2, 32 x 512^2, 24 x 1024^2, 8 x 2048^2; PSF twice
  the image, FWHM U(3, 8) px; moresane S=5, gain 0.2, 100 major / 30 minor iterations) run one at a
  time and batched (size classes in lockstep, classes concurrent): every job's model and residual
  identical bit for bit; jobs 0 (256^2) and 32 (512^2) against the float64 oracle's runs of them
  (tests/golden/c3_oracle.npz, written by tests/golden/make_c3_fixtures.py): identical decision trace
  and model within 1e-4 in the float32 mode (the residual within what the model's distance implies),
  model and residual within 1e-4 in the float64 mode.
* Served mix: the Poisson request stream (60 requests at 20/s, sides 512/1024/2048/4096 at
  50/30/15/5 %, S = log2(side) - 5) through the continuous-batching server and through the plain
  one-at-a-time server: every request's result identical bit for bit.
"""

import json
import os
import sys
import time

import numpy as np
import pytest

from pymoresane_amd import batch, synth

pytestmark = pytest.mark.gpu

KW = dict(loop_gain=0.2, sigma_level=4, major_loop_miter=100, tolerance=0.75, enforce_positivity=True,
          conv_mode="linear", minor_loop_miter=30, accuracy=1e-6)
C3_FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c3_oracle.npz")


def say(msg):
    print(msg)
    sys.stdout.flush()


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def discrete(trace):
    return [tuple(v for v in e if not isinstance(v, float)) for e in trace]


@pytest.fixture(scope="module")
def c3_jobs(hip):
    sides = synth.config_c3_sides()
    say("c3: generating %d pairs" % len(sides))
    pairs = [synth.config_c3_pair(i, s) for i, s in enumerate(sides)]
    jobs = batch.DeviceJobs(pairs, [dict(KW, scale_count=5)] * len(pairs))
    return sides, pairs, jobs


def test_ragged_batch_equals_one_at_a_time(c3_jobs):
    sides, _, jobs = c3_jobs
    order = sorted(range(len(sides)), key=lambda i: -sides[i])
    t0 = time.time()
    batch.run_batch(jobs, n_workers=1, max_batch=1, order=order)
    say("c3 plain: %.2f s" % (time.time() - t0))
    ref = jobs.results()
    jobs.reset()
    t0 = time.time()
    stats = batch.run_batch(jobs, n_workers=8, max_batch=32, order=order)
    say("c3 batched: %.2f s" % (time.time() - t0))
    assert all(s[0] == 0 for s in stats)
    for i, ((m, r), (m0, r0)) in enumerate(zip(jobs.results(), ref)):
        assert np.array_equal(m, m0) and np.array_equal(r, r0), "job %d (%d^2) differs" % (i, sides[i])


def test_ragged_batch_jobs_match_the_oracle(c3_jobs):
    from scipy import signal
    sides, pairs, jobs = c3_jobs
    fx = np.load(C3_FIXTURE)
    assert json.loads(str(fx["kw"])) == dict(KW, scale_count=5)
    results = jobs.results()  # the batched run of the previous test
    for i in fx["jobs"]:
        dirty, psf = pairs[i]
        n = dirty.shape[0]
        say("c3 job %d (%d^2) vs the oracle" % (i, n))
        model = np.zeros((n, n))
        model.flat[fx["j%d_model_idx" % i]] = fx["j%d_model_val" % i]
        resid = dirty.astype(np.float64) - signal.fftconvolve(model, psf.astype(np.float64))[n:2 * n, n:2 * n]
        assert abs(np.linalg.norm(resid) / float(fx["j%d_residual_norm" % i]) - 1) < 1e-10
        np.testing.assert_allclose(resid[::32], fx["j%d_residual_rows" % i], rtol=0, atol=1e-12)
        m, r = results[i]
        # the float32 residual against dirty - H m of its own model, computed here in float64:
        # isolates the arithmetic of the residual from the model's distance to the oracle's
        own = dirty.astype(np.float64) - signal.fftconvolve(m.astype(np.float64), psf.astype(np.float64))[n:2 * n,
                                                                                                        n:2 * n]
        gain = np.linalg.norm(dirty - resid) / np.linalg.norm(resid)
        say("  float32: model %.3g, residual %.3g (vs its own model %.3g), ||H m|| / ||r|| = %.1f"
            % (rel(m, model), rel(r, resid), rel(r, own), gain))
        # the bar on the deconvolved model; the residual is dirty - H m to float32 arithmetic, so its
        # distance to the oracle's is the model's carried through H and divided by the small
        # residual (||H m|| / ||r|| is 92 for job 0): the float64 mode below meets 1e-4 there too
        assert rel(m, model) < 1e-4, rel(m, model)
        assert rel(r, own) < 1e-4, rel(r, own)
        assert rel(r, resid) < 2 * gain * rel(m, model) + rel(r, own)
    # the decision trace of each checked job, from a plain run of it (the batch records no per-job
    # trace), and the same job in the float64 mode, model and residual within 1e-4
    from pymoresane_amd.main import DataImage
    for i in fx["jobs"]:
        dirty, psf = pairs[i]
        n = dirty.shape[0]
        ref_trace = [tuple(e) for e in json.loads(str(fx["j%d_trace" % i]))]
        img = DataImage(dirty, psf, None, 1.0, 1.0)
        img.moresane(**dict(KW, scale_count=5))
        assert discrete(img.trace) == discrete(ref_trace)
        np.testing.assert_array_equal(img.model, results[i][0])
        img = DataImage(dirty, psf, None, 1.0, 1.0)
        img.moresane(precision="float64", **dict(KW, scale_count=5))
        assert discrete(img.trace) == discrete(ref_trace)
        model = np.zeros((n, n))
        model.flat[fx["j%d_model_idx" % i]] = fx["j%d_model_val" % i]
        resid = dirty.astype(np.float64) - signal.fftconvolve(model, psf.astype(np.float64))[n:2 * n, n:2 * n]
        say("  float64 job %d: model %.3g, residual %.3g" % (i, rel(img.model, model), rel(img.residual, resid)))
        assert rel(img.model, model) < 1e-4 and rel(img.residual, resid) < 1e-4


def test_served_request_stream_equals_one_at_a_time(hip):
    reqs = synth.config_c4_requests(60, rate=20.0)
    say("c4: generating %d requests" % len(reqs))
    psfs = {}
    inputs = []
    for t_arr, side, seed in reqs:
        rng = np.random.default_rng(seed)
        if side not in psfs:
            psfs[side] = synth.make_psf(2 * side)
        sky = synth.make_sky(side, rng)
        dirty = (synth.linear_convolve(sky, psfs[side]) + rng.normal(0.0, 1e-3, (side, side))).astype(np.float32)
        inputs.append((t_arr, side, dirty))
    results = {}
    for label, mb in (("plain", 1), ("served", 32)):
        srv = batch.Server(max_batch=mb, budget_bytes=200e9)
        handles = []
        t0 = time.perf_counter()
        for t_arr, side, dirty in inputs:
            wait = t_arr - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            handles.append(srv.submit(dirty, psfs[side], scale_count=int(np.log2(side)) - 5, **KW))
        srv.drain()
        say("c4 %s: %.2f s" % (label, time.perf_counter() - t0))
        assert all(h["req"].status == 0 for h in handles)
        results[label] = [(h["model"], h["resid"]) for h in handles]
        srv.close()
    for k, ((m, r), (m0, r0)) in enumerate(zip(results["served"], results["plain"])):
        assert np.array_equal(m, m0) and np.array_equal(r, r0), "request %d (%d^2) differs" % (k, inputs[k][1])
