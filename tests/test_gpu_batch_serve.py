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


def pairs():
    out = []
    for i, n in enumerate([64, 128, 64, 256, 128, 64]):
        out.append(synth.make_pair(n, 500 + i, n_points=6 + i, n_blobs=1))
    return out


def alone(pairs_):
    res = []
    for d, p in pairs_:
        img = DataImage(d, p, None, 1.0, 1.0)
        img.moresane(**KW)
        res.append((img.model, img.residual))
    return res


def test_batch_equals_alone(hip):
    say("batch vs alone, 6 ragged jobs, 4 workers")
    ps = pairs()
    ref = alone(ps)
    jobs = batch.DeviceJobs(ps, [KW] * len(ps))
    stats = batch.run_batch(jobs, n_workers=4)
    assert all(s[0] == 0 for s in stats)
    for (m, r), (m0, r0) in zip(jobs.results(), ref):
        np.testing.assert_array_equal(m, m0)
        np.testing.assert_array_equal(r, r0)
    # a second run on the same (reset) device state gives the same bits again
    jobs.reset()
    batch.run_batch(jobs, n_workers=3, order=[5, 4, 3, 2, 1, 0])
    for (m, r), (m0, r0) in zip(jobs.results(), ref):
        np.testing.assert_array_equal(m, m0)
        np.testing.assert_array_equal(r, r0)


def test_served_mix_equals_alone(hip):
    say("server: staggered arrivals, 3 workers")
    ps = pairs()
    ref = alone(ps)
    srv = batch.Server(n_workers=3, budget_bytes=8e9)
    reqs = []
    for d, p in ps:
        reqs.append(srv.submit(d, p, **KW))
        time.sleep(0.002)
    srv.drain()
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
    probe = batch.Server(n_workers=1, budget_bytes=64e9)
    fp = [probe.submit(d, p, **KW) for d, p in ps]
    probe.drain()
    need = max(rq["req"].footprint for rq in fp)
    probe.close()
    srv = batch.Server(n_workers=3, budget_bytes=float(need))
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
    srv = batch.Server(n_workers=1, budget_bytes=1e6)
    d, p = synth.make_pair(128, 1, n_points=3, n_blobs=0)
    with pytest.raises(Exception, match="budget"):
        srv.submit(d, p, **KW)
    srv.close()
