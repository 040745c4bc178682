"""Batched and served MORESANE on one MI355X.

* ``run_batch`` -- many independent dirty/PSF pairs (ragged sizes), inputs
  resident in HBM: jobs of one size class are stepped in lockstep, so every
  kernel of a CG iteration is one launch for the whole batch; classes run
  concurrently on their own streams.
* ``Server``    -- continuous batching: requests queue per size class and join
  the class's running batch between steps as slots and HBM budget free up.
Both produce, per job, exactly the result of running that job alone.
"""

import ctypes

import numpy as np

from . import _lib
from ._lib import BatchJob, DeviceArray, ServeRequest, check
from .main import make_params

DEFAULTS = dict(subregion=None, scale_count=None, sigma_level=4, loop_gain=0.1, tolerance=0.75, accuracy=1e-6,
                major_loop_miter=100, minor_loop_miter=30, conv_mode="linear", enforce_positivity=False,
                edge_suppression=False, edge_offset=0, flux_threshold=0, neg_comp=False, edge_excl=0, int_excl=0,
                precision="float32")


def params_for(n, psf_n, **kw):
    k = dict(DEFAULTS)
    unknown = set(kw) - set(k)
    if unknown:
        raise TypeError("unknown MORESANE arguments: %s" % sorted(unknown))
    k.update(kw)
    return make_params(n, psf_n, k["subregion"], k["scale_count"], k["sigma_level"], k["loop_gain"], k["tolerance"],
                       k["accuracy"], k["major_loop_miter"], k["minor_loop_miter"], k["conv_mode"],
                       k["enforce_positivity"], k["edge_suppression"], k["edge_offset"], k["flux_threshold"],
                       k["neg_comp"], k["edge_excl"], k["int_excl"], k["precision"])


class DeviceJobs(object):
    """Inputs/outputs of a list of jobs resident in HBM."""

    def __init__(self, pairs, kwargs_list):
        self.items = []
        for (dirty, psf), kw in zip(pairs, kwargs_list):
            dirty = np.ascontiguousarray(dirty, dtype=np.float32)
            psf = np.ascontiguousarray(psf, dtype=np.float32)
            d = DeviceArray.from_host(dirty)
            self.items.append(dict(dirty=d, psf=DeviceArray.from_host(psf), model=DeviceArray(dirty.shape),
                                   resid=DeviceArray(dirty.shape), prm=params_for(dirty.shape[0], psf.shape[0], **kw),
                                   kw=kw))
        self.reset()

    def reset(self):
        for it in self.items:
            it["model"].fill(0.0)
            check(_lib.lib().pms_memcpy_d2d(it["resid"].ptr, it["dirty"].ptr, it["dirty"].nbytes, None))
        # the fills and copies above run on the null stream, which the runners' non-blocking streams
        # do not wait for
        _lib.synchronize()

    def results(self):
        return [(it["model"].get(), it["resid"].get()) for it in self.items]


def run_batch(jobs, n_workers=4, max_batch=32, order=None, by_scale=False, start_scale=1, stop_scale=20):
    """Run every job of a DeviceJobs set; returns per-job (status, major_iters, seconds, t_start, wait_seconds).

    Jobs are grouped by size class; up to ``n_workers`` classes run at once, each stepping up to
    ``max_batch`` jobs in lockstep with batched launches.  ``n_workers=1, max_batch=1`` is the
    plain one-job-at-a-time path.
    """
    lib = _lib.lib()
    idx = list(range(len(jobs.items))) if order is None else list(order)
    arr = (BatchJob * len(idx))()
    for slot, i in enumerate(idx):
        it = jobs.items[i]
        b = arr[slot]
        b.dirty, b.psf, b.mask = it["dirty"].ptr, it["psf"].ptr, None
        b.model, b.residual = it["model"].ptr, it["resid"].ptr
        b.prm = it["prm"]
        b.by_scale, b.start_scale, b.stop_scale = int(by_scale), int(start_scale), int(stop_scale)
    check(lib.pms_batch_run(arr, len(idx), int(n_workers), int(max_batch)))
    out = [None] * len(jobs.items)
    for slot, i in enumerate(idx):
        b = arr[slot]
        if b.status != 0:
            raise _lib.PmsError(b.status, "batch job %d failed" % i)
        out[i] = (b.status, b.major_iters, b.seconds, b.t_start, b.wait_seconds)
    return out


class Server(object):
    """Continuous-batching MORESANE server: one lockstep batch of up to ``max_batch`` requests per
    size class, requests admitted between steps under a shared HBM budget."""

    def __init__(self, max_batch=32, budget_bytes=200e9):
        self._lib = _lib.lib()
        self._h = ctypes.c_void_p(0)
        check(self._lib.pms_server_create(int(max_batch), float(budget_bytes), ctypes.byref(self._h)))
        self._live = []

    def now(self):
        t = ctypes.c_double(0.0)
        check(self._lib.pms_server_now(self._h, ctypes.byref(t)))
        return t.value

    def submit(self, dirty, psf, mask=None, **kw):
        dirty = np.ascontiguousarray(dirty, dtype=np.float32)
        psf = np.ascontiguousarray(psf, dtype=np.float32)
        req = ServeRequest()
        keep = dict(dirty=dirty, psf=psf, model=np.zeros_like(dirty), resid=np.zeros_like(dirty), req=req)
        if mask is not None:
            keep["mask"] = np.ascontiguousarray(mask, dtype=np.float32)
            req.mask_host = keep["mask"].ctypes.data
        req.dirty_host = dirty.ctypes.data
        req.psf_host = psf.ctypes.data
        req.model_host = keep["model"].ctypes.data
        req.residual_host = keep["resid"].ctypes.data
        req.prm = params_for(dirty.shape[0], psf.shape[0], **kw)
        check(self._lib.pms_server_submit(self._h, ctypes.byref(req)))
        self._live.append(keep)
        return keep

    def drain(self):
        """Wait until every submitted request has finished; the server then drops its references to
        their host buffers (the caller's dicts from submit() keep the results)."""
        check(self._lib.pms_server_drain(self._h))
        self._live = [k for k in self._live if k["req"].status > 0]  # > 0: queued or running

    def close(self):
        if self._h.value:
            self._lib.pms_server_destroy(self._h)
            self._h = ctypes.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
