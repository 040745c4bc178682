"""IUWT (isotropic undecimated wavelet transform) on the MI355X.

Drop-in for pymoresane/iuwt.py:17-64: same function names, positional order
and keyword defaults apart from ``mode``, whose HIP value is ``'hip'`` (the
reference's ``'gpu'`` is accepted as an alias).  The reference's CPU modes
``'ser'`` and ``'mp'`` are not part of this package, which only provides the
HIP path; asking for them, or for an unknown mode, raises ``ValueError``
(the reference silently returns ``None`` for unknown modes, iuwt.py:36-41).
"""

import numpy as np

from . import _lib
from ._lib import DeviceArray, Workspace, as_device, check

HIP_MODES = ("hip", "gpu")


def check_mode(mode, what="mode"):
    if mode in HIP_MODES:
        return
    if mode in ("ser", "mp", "cpu"):
        raise ValueError("%s=%r is a CPU path of the reference; pymoresane_amd provides the HIP path "
                         "('hip', or 'gpu' as an alias)" % (what, mode))
    raise ValueError("unknown %s %r: expected 'hip' (or 'gpu')" % (what, mode))


def iuwt_decomposition(in1, scale_count, scale_adjust=0, mode="hip", core_count=2, store_smoothed=False,
                       store_on_gpu=False):
    """Detail coefficients of scales scale_adjust..scale_count-1 (pymoresane/iuwt.py:17-41).

    ``in1`` is a 2-D numpy array or DeviceArray.  Returns float32 numpy
    arrays, or DeviceArrays with ``store_on_gpu``; with ``store_smoothed`` the
    smoothest approximation is returned as well (also with ``store_on_gpu``,
    where the reference drops it, iuwt.py:494-495).  ``core_count`` only
    applies to the reference's 'mp' mode and is ignored.
    """
    check_mode(mode)
    src, _ = as_device(in1)
    if src.ndim != 2:
        raise ValueError("iuwt_decomposition expects a 2-D image, got shape %s" % (src.shape,))
    h, w = src.shape
    scale_count, scale_adjust = int(scale_count), int(scale_adjust)
    if scale_count <= scale_adjust or scale_adjust < 0:
        raise ValueError("need scale_count > scale_adjust >= 0")
    lib = _lib.lib()
    detail = DeviceArray((scale_count - scale_adjust, h, w))
    smooth = DeviceArray((h, w)) if store_smoothed else None
    work = Workspace.get(lib.pms_iuwt_work_bytes(h, w))
    check(lib.pms_iuwt_decompose(src.ptr, h, w, scale_count, scale_adjust, detail.ptr,
                                 smooth.ptr if smooth is not None else None, work.ptr, None))
    if store_on_gpu:
        return (detail, smooth) if store_smoothed else detail
    if store_smoothed:
        return detail.get(), smooth.get()
    return detail.get()


def iuwt_recomposition(in1, scale_adjust=0, mode="hip", core_count=1, store_on_gpu=False, smoothed_array=None):
    """Inverse transform of ``in1`` (planes of scales scale_adjust..) (pymoresane/iuwt.py:43-64)."""
    check_mode(mode)
    det, _ = as_device(in1)
    if det.ndim != 3:
        raise ValueError("iuwt_recomposition expects (scales, rows, cols), got shape %s" % (det.shape,))
    n, h, w = det.shape
    lib = _lib.lib()
    smooth = None
    if smoothed_array is not None:
        smooth, _ = as_device(smoothed_array)
        if smooth.shape != (h, w):
            raise ValueError("smoothed_array shape %s does not match planes %s" % (smooth.shape, (h, w)))
    out = DeviceArray((h, w))
    work = Workspace.get(lib.pms_iuwt_work_bytes(h, w))
    check(lib.pms_iuwt_recompose(det.ptr, n, h, w, int(scale_adjust), smooth.ptr if smooth is not None else None,
                                 out.ptr, work.ptr, None))
    return out if store_on_gpu else out.get()
