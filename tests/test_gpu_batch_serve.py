"""Batched and served forms: every job's result equals running it alone, bit for bit."""

import sys
import time

import numpy as np
import pytest

from pymoresane_amd import batch, synth
from pymoresane_amd.main import DataImage

pytestmark = pytest.mark.gpu

KW = dict(loop_gain=0.2, major_loop_miter=12, enforce_positivity=True, scale_count=3)


def say(msg):
    print(msg)
    sys.stdout.flush()


SIDES = [64, 128, 64, 256, 128, 64, 1024, 1024]


def pairs(sides=SIDES):
    out = []
    for i, n in enumerate(sides):
        out.append(synth.make_pair(n, 500 + i, n_points=6 + i, n_blobs=1))
    return out


def alone(pairs_, by_scale=False):
    res = []
    for d, p in pairs_:
        img = DataImage(d, p, None, 1.0, 1.0)
        if by_scale:
            img.moresane_by_scale(start_scale=1, stop_scale=3, **BYSCALE_KW)
        else:
            img.moresane(**KW)
        res.append((img.model, img.residual))
    return res


BYSCALE_KW = dict(loop_gain=0.2, major_loop_miter=6, enforce_positivity=True)


def test_batch_equals_alone(hip):
    say("batch vs alone, 8 ragged jobs in 4 size classes, lockstep batches")
    ps = pairs()
    ref = alone(ps)
    jobs = batch.DeviceJobs(ps, [KW] * len(ps))
    stats = batch.run_batch(jobs, n_workers=4, max_batch=32)
    assert all(s[0] == 0 for s in stats)
    for (m, r), (m0, r0) in zip(jobs.results(), ref):
        np.testing.assert_array_equal(m, m0)
        np.testing.assert_array_equal(r, r0)
    # fewer slots than jobs per class (slots refill as jobs finish), one worker for every class
    jobs.reset()
    batch.run_batch(jobs, n_workers=1, max_batch=2, order=[7, 5, 4, 3, 2, 1, 0, 6])
    for (m, r), (m0, r0) in zip(jobs.results(), ref):
        np.testing.assert_array_equal(m, m0)
        np.testing.assert_array_equal(r, r0)
    # plain path through the same entry point
    jobs.reset()
    batch.run_batch(jobs, n_workers=1, max_batch=1)
    for (m, r), (m0, r0) in zip(jobs.results(), ref):
        np.testing.assert_array_equal(m, m0)
        np.testing.assert_array_equal(r, r0)


def test_batch_by_scale_equals_alone(hip):
    say("batch vs alone, by-scale jobs (scale counts differ across the batch)")
    ps = pairs([64, 64, 128, 64])
    ref = alone(ps, by_scale=True)
    jobs = batch.DeviceJobs(ps, [BYSCALE_KW] * len(ps))
    stats = batch.run_batch(jobs, n_workers=2, max_batch=8, by_scale=True, start_scale=1, stop_scale=3)
    assert all(s[0] == 0 for s in stats)
    for (m, r), (m0, r0) in zip(jobs.results(), ref):
        np.testing.assert_array_equal(m, m0)
        np.testing.assert_array_equal(r, r0)


def test_served_mix_equals_alone(hip):
    say("server: staggered arrivals, batches of up to 3 per size class")
    ps = pairs()
    ref = alone(ps)
    srv = batch.Server(max_batch=3, budget_bytes=16e9)
    reqs = []
    for d, p in ps:
        reqs.append(srv.submit(d, p, **KW))
        time.sleep(0.002)
    srv.drain()
    assert srv._live == []  # finished requests are released; the caller's dicts keep the results
    for rq, (m0, r0) in zip(reqs, ref):
        assert rq["req"].status == 0
        np.testing.assert_array_equal(rq["model"], m0)
        np.testing.assert_array_equal(rq["resid"], r0)
    starts = sorted((rq["req"].t_start, rq["req"].t_done) for rq in reqs)
    srv.close()
    # with room for everything the jobs overlapped on the device
    assert any(starts[i + 1][0] < starts[i][1] for i in range(len(starts) - 1))


def test_server_budget_serialises(hip):
    say("server: budget admits one request at a time")
    ps = pairs()[:3]
    ref = alone(ps)
    probe = batch.Server(max_batch=1, budget_bytes=64e9)
    fp = [probe.submit(d, p, **KW) for d, p in ps]
    probe.drain()
    need = max(rq["req"].footprint for rq in fp)
    probe.close()
    srv = batch.Server(max_batch=3, budget_bytes=float(need))
    reqs = [srv.submit(d, p, **KW) for d, p in ps]
    srv.drain()
    spans = sorted((rq["req"].t_start, rq["req"].t_done) for rq in reqs)
    srv.close()
    for i in range(len(spans) - 1):
        assert spans[i + 1][0] >= spans[i][1]
    for rq, (m0, r0) in zip(reqs, ref):
        np.testing.assert_array_equal(rq["model"], m0)
        np.testing.assert_array_equal(rq["resid"], r0)


def test_server_rejects_oversized_request(hip):
    srv = batch.Server(max_batch=1, budget_bytes=1e6)
    d, p = synth.make_pair(128, 1, n_points=3, n_blobs=0)
    with pytest.raises(Exception, match="budget"):
        srv.submit(d, p, **KW)
    srv.close()


def test_batch_of_4096_frames_equals_alone(hip):
    """Two 4096^2 jobs with 8192^2 PSFs (linear mode) in one batch: the batched launches at the 8192
    frame, the largest of the served mix, give every job the bits it gets alone."""
    say("batch vs alone, 8192-frame convolutions, two jobs per batch")
    kw = dict(loop_gain=0.2, major_loop_miter=2, enforce_positivity=True, scale_count=3)
    ps = [synth.make_pair(4096, 700 + i, n_points=8, n_blobs=1) for i in range(2)]
    ref = []
    for d, p in ps:
        img = DataImage(d, p, None, 1.0, 1.0)
        img.moresane(**kw)
        ref.append((img.model, img.residual))
    jobs = batch.DeviceJobs(ps, [kw] * len(ps))
    stats = batch.run_batch(jobs, n_workers=1, max_batch=2)
    assert all(s[0] == 0 for s in stats)
    for (m, r), (m0, r0) in zip(jobs.results(), ref):
        np.testing.assert_array_equal(m, m0)
        np.testing.assert_array_equal(r, r0)
