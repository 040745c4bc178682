"""Noise estimate, thresholding, source extraction and SNR on the MI355X.

Drop-in for pymoresane/iuwt_toolbox.py:17-306.  ``estimate_threshold`` and
``apply_threshold`` have no mode argument in the reference; here they take an
optional trailing ``mode='hip'``.  ``source_extraction`` keeps its signature
with ``mode='hip'`` (``'gpu'`` accepted); ``snr_ratio`` takes a trailing
``mode='hip'``.  Inputs may be numpy arrays or DeviceArrays.
"""

import ctypes

import numpy as np

from . import _lib
from ._lib import DeviceArray, Workspace, as_device, check
from .iuwt import check_mode


def _stack(in1):
    src, _ = as_device(in1)
    if src.ndim == 2:
        return src, (1,) + src.shape, True
    if src.ndim != 3:
        raise ValueError("expected a 2-D plane or a (scales, rows, cols) stack, got %s" % (src.shape,))
    return src, src.shape, False


def estimate_threshold(in1, edge_excl=0, int_excl=0, mode="hip"):
    """Per-scale MAD noise estimate median(|w|)/0.6745 (pymoresane/iuwt_toolbox.py:17-48), float64 numpy."""
    check_mode(mode)
    src, (s, h, w), _ = _stack(in1)
    lib = _lib.lib()
    thr = DeviceArray((s,), np.float64)
    work = Workspace.get(lib.pms_threshold_work_bytes(s, h, w))
    check(lib.pms_estimate_threshold(src.ptr, s, h, w, int(edge_excl), int(int_excl), thr.ptr, work.ptr, None))
    return thr.get()


def apply_threshold(in1, threshold, sigma_level=4, mode="hip"):
    """Keep coefficients with |w| > sigma_level * threshold (pymoresane/iuwt_toolbox.py:50-75)."""
    check_mode(mode)
    on_dev = isinstance(in1, DeviceArray)
    src, (s, h, w), flat = _stack(in1)
    thr = np.atleast_1d(np.asarray(threshold, dtype=np.float64))
    if thr.size == 1 and s > 1:
        raise ValueError("one threshold per scale is required for a plane stack")
    lib = _lib.lib()
    thr_dev = DeviceArray.from_host(thr[:s] if thr.size >= s else thr)
    out = DeviceArray(src.shape)
    work = Workspace.get(lib.pms_threshold_work_bytes(s, h, w))
    check(lib.pms_apply_threshold(src.ptr, s, h, w, thr_dev.ptr, float(sigma_level), out.ptr, work.ptr, None))
    return out if on_dev else out.get()


def source_extraction(in1, tolerance, mode="hip", store_on_gpu=False, neg_comp=False, return_labels=False):
    """Extract significant objects coarse to fine (pymoresane/iuwt_toolbox.py:77-154).

    Returns ``(objects * in1, objects)``: float32 extracted coefficients and the
    int32 {0,1} mask (DeviceArrays with ``store_on_gpu``, the mask then being
    uint8).  ``return_labels`` adds the per-scale object labels (int32, raster
    order of each object's first pixel, as ``scipy.ndimage.label`` numbers them).
    """
    check_mode(mode)
    src, (s, h, w), flat = _stack(in1)
    lib = _lib.lib()
    ext = DeviceArray((s, h, w))
    mask = DeviceArray((s, h, w), np.uint8)
    labels = DeviceArray((s, h, w), np.int32) if return_labels else None
    work = Workspace.get(lib.pms_extract_work_bytes(s, h, w))
    check(lib.pms_source_extraction(src.ptr, s, h, w, float(tolerance), int(bool(neg_comp)), ext.ptr, mask.ptr,
                                    labels.ptr if labels is not None else None, work.ptr, None))
    if store_on_gpu:
        res = (ext, mask)
    else:
        res = (ext.get().reshape(src.shape), mask.get().astype(np.int32).reshape(src.shape))
    if return_labels:
        res = res + ((labels if store_on_gpu else labels.get().reshape(src.shape)),)
    return res


def snr_ratio(in1, in2, mode="hip"):
    """20 log10(||in1|| / ||in1 - in2||) with float64 sums (pymoresane/iuwt_toolbox.py:292-306)."""
    check_mode(mode)
    a, _ = as_device(in1)
    b, _ = as_device(in2)
    if a.size != b.size:
        raise ValueError("size mismatch %s vs %s" % (a.shape, b.shape))
    lib = _lib.lib()
    work = Workspace.get(lib.pms_snr_work_bytes(a.size))
    out = ctypes.c_double(0.0)
    check(lib.pms_snr_ratio(a.ptr, b.ptr, a.size, ctypes.byref(out), work.ptr, None))
    return out.value
