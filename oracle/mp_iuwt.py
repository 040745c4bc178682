"""Restatement of the reference multi-process IUWT, for CPU timing only.

Follows pymoresane/iuwt.py:193-382 (``mp_iuwt_decomposition``,
``mp_iuwt_recomposition``, ``mp_a_trous``, ``mp_a_trous_kernel``): every
a trous call copies the image into a shared float32 array and forks
``core_count`` processes twice -- once per axis -- each filtering one strip
in place.  Unlike the reference (iuwt.py:310, 321) the last strip takes the
remainder rows/columns instead of dropping them.  Test infrastructure: it is
timed as the ``mp`` CPU baseline, never used as a checker of record.
"""

import ctypes
import multiprocessing as mp

import numpy as np

from .iuwt64 import B3, check_scale, reflect_index


def _strip_pass(shared, shape, scale, lo, hi, axis):
    arr = np.frombuffer(shared, dtype=np.float32).reshape(shape)
    d = 1 << scale
    n = shape[axis]
    idx = np.arange(n)
    if axis == 0:
        src = arr[:, lo:hi].astype(np.float64)
        acc = np.zeros_like(src)
        for k in range(-2, 3):
            acc += B3[k + 2] * src[reflect_index(idx + k * d, n), :]
        arr[:, lo:hi] = acc
    else:
        src = arr[lo:hi, :].astype(np.float64)
        acc = np.zeros_like(src)
        for k in range(-2, 3):
            acc += B3[k + 2] * src[:, reflect_index(idx + k * d, n)]
        arr[lo:hi, :] = acc


def mp_a_trous(c0, scale, core_count):
    c0 = np.asarray(c0)
    check_scale(c0.shape, scale)
    shape = c0.shape
    shared = mp.RawArray(ctypes.c_float, int(np.prod(shape)))
    np.frombuffer(shared, dtype=np.float32).reshape(shape)[:] = c0
    ctx = mp.get_context("fork")
    for axis, n in ((0, shape[1]), (1, shape[0])):
        width = -(-n // core_count)
        procs = []
        for i in range(core_count):
            lo, hi = i * width, min((i + 1) * width, n)
            if lo >= hi:
                continue
            pr = ctx.Process(target=_strip_pass, args=(shared, shape, scale, lo, hi, axis))
            pr.start()
            procs.append(pr)
        for pr in procs:
            pr.join()
    return np.frombuffer(shared, dtype=np.float32).reshape(shape).copy()


def mp_iuwt_decomposition(in1, scale_count, scale_adjust=0, store_smoothed=False, core_count=2):
    c0 = np.asarray(in1, dtype=np.float32)
    for s in range(scale_adjust):
        c0 = mp_a_trous(c0, s, core_count)
    detail = np.empty((scale_count - scale_adjust,) + c0.shape)
    for s in range(scale_adjust, scale_count):
        c = mp_a_trous(c0, s, core_count)
        detail[s - scale_adjust] = c0 - mp_a_trous(c, s, core_count)
        c0 = c
    if store_smoothed:
        return detail, c0
    return detail


def mp_iuwt_recomposition(in1, scale_adjust=0, core_count=1, smoothed_array=None):
    top = in1.shape[0] + scale_adjust
    rec = np.zeros(in1.shape[1:], dtype=np.float32) if smoothed_array is None else smoothed_array
    for s in range(top - 1, scale_adjust - 1, -1):
        rec = mp_a_trous(rec, s, core_count) + in1[s - scale_adjust]
    for s in range(scale_adjust - 1, -1, -1):
        rec = mp_a_trous(rec, s, core_count)
    return rec
