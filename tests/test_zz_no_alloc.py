"""The allocation-free median entry point (pms_estimate_threshold_noalloc) gives the same
thresholds, bit for bit, as pms_estimate_threshold with a caller workspace, through the C ABI,
on two calls in a row (the second smaller, with a noise mask), and rejects a stack larger than
its fixed device array."""

import ctypes

import numpy as np
import pytest

from oracle import toolbox64 as otools
from pymoresane_amd import _lib
from pymoresane_amd._lib import DeviceArray, PmsError, check

pytestmark = pytest.mark.gpu


def thresholds(lib, planes, edge_excl, int_excl, noalloc):
    s, h, w = planes.shape
    d = DeviceArray.from_host(planes)
    out = DeviceArray((s,), dtype=np.float64)
    if noalloc:
        check(lib.pms_estimate_threshold_noalloc(d.ptr, s, h, w, edge_excl, int_excl, out.ptr, None))
    else:
        work = DeviceArray((lib.pms_threshold_work_bytes(s, h, w),), dtype=np.uint8)
        check(lib.pms_estimate_threshold(d.ptr, s, h, w, edge_excl, int_excl, out.ptr, work.ptr, None))
    return out.get()


def test_noalloc_thresholds_equal_the_workspace_entry_point(hip):
    rng = np.random.default_rng(21)
    cases = [(rng.standard_normal((7, 96, 80)).astype(np.float32), 0, 0),   # full-size stack first
             (rng.standard_normal((3, 33, 47)).astype(np.float32), 3, 5)]   # then a smaller, masked one
    for planes, ee, ie in cases:
        print("case", planes.shape, ee, ie, flush=True)
        got = thresholds(hip, planes, ee, ie, noalloc=True)
        ref = thresholds(hip, planes, ee, ie, noalloc=False)
        np.testing.assert_array_equal(got, ref)
        np.testing.assert_array_equal(got, otools.estimate_threshold(planes, ee, ie))


def test_noalloc_rejects_more_planes_than_its_array(hip):
    planes = DeviceArray((17, 8, 8))
    out = DeviceArray((17,), dtype=np.float64)
    st = hip.pms_estimate_threshold_noalloc(planes.ptr, 17, 8, 8, 0, 0, out.ptr, None)
    assert st == -1
    assert b"1..16 planes" in hip.pms_last_error()
