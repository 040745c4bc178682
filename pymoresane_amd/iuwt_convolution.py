"""FFT convolution with a precomputed PSF spectrum on the MI355X.

Drop-in for pymoresane/iuwt_convolution.py:18-158.  ``fft_convolve(in1, in2,
conv_device, conv_mode, store_on_gpu)`` keeps the reference's signature; its
HIP device value is ``'hip'`` (``'gpu'`` is an alias).  ``in2`` is either a
``PsfSpectrum`` built by ``psf_spectrum`` / ``gpu_r2c_fft(...,
store_on_gpu=True)``, or a numpy ``rfft2`` spectrum as the reference's CPU
path takes (converted to the library layout on each call).
"""

import ctypes

import numpy as np

from . import _lib
from ._lib import DeviceArray, Workspace, as_device, check
from .iuwt import check_mode


class PsfSpectrum(object):
    """A PSF spectrum resident in HBM, for a p x p convolution frame.

    Layout (library-internal): p/2+1 frequency columns of p complex64 values
    each, in the library's FFT storage order, carrying the p/2 roll that puts
    the PSF centre at the frame origin.
    """

    def __init__(self, p, buf):
        self.p = int(p)
        self.buf = buf

    @property
    def ptr(self):
        return self.buf.ptr

    @classmethod
    def from_rfft2(cls, spec):
        """Convert a numpy ``rfft2`` spectrum (p x (p/2+1)) of a PSF frame to the device layout."""
        spec = np.asarray(spec)
        p = spec.shape[0]
        if spec.ndim != 2 or spec.shape[1] != p // 2 + 1:
            raise ValueError("expected a p x (p/2+1) rfft2 spectrum, got shape %s" % (spec.shape,))
        lib = _lib.load()
        host_in = np.ascontiguousarray(spec.astype(np.complex64))
        nbytes = lib.pms_fft_spectrum_bytes(p)
        host_out = np.empty(nbytes // 8, dtype=np.complex64)
        check(lib.pms_spectrum_from_rfft2(host_in.ctypes.data, p, host_out.ctypes.data))
        return cls(p, DeviceArray.from_host(host_out))

    def to_rfft2(self):
        """The spectrum in numpy ``rfft2`` layout (host, complex64)."""
        lib = _lib.load()
        p = self.p
        raw = self.buf.get().view(np.complex64).reshape(p // 2 + 1, p)
        # invert the storage permutation: probe positions via a unit-impulse conversion is costly; use the index map
        pos = spectrum_positions(p)
        out = raw[:, pos].T.copy()
        m = np.arange(p)[:, None]
        k = np.arange(p // 2 + 1)[None, :]
        out *= np.where((m + k) % 2 == 1, -1, 1).astype(np.complex64)
        return out


_POS_CACHE = {}


def spectrum_positions(p):
    """Storage position of each frequency index 0..p-1 along a spectrum column."""
    if p not in _POS_CACHE:
        lib = _lib.load()
        # the conversion routine writes frequency m to its position: feed m as the value and read it back
        probe = np.zeros((p, p // 2 + 1), dtype=np.complex64)
        probe[:, 0] = np.arange(p) * np.where(np.arange(p) % 2 == 1, -1, 1)
        out = np.empty((p // 2 + 1) * p, dtype=np.complex64)
        check(lib.pms_spectrum_from_rfft2(probe.ctypes.data, p, out.ctypes.data))
        col0 = out[:p].real.astype(np.int64)
        pos = np.empty(p, dtype=np.int64)
        pos[col0] = np.arange(p)
        _POS_CACHE[p] = pos
    return _POS_CACHE[p]


def pad_array(in1):
    """Zero frame of twice the size with in1 in the centre (pymoresane/iuwt_convolution.py:142-158), host."""
    in1 = np.asarray(in1)
    h, w = in1.shape
    out = np.zeros((2 * h, 2 * w), dtype=in1.dtype)
    out[h // 2:h // 2 + h, w // 2:w // 2 + w] = in1
    return out


def psf_spectrum(psf, p=None):
    """Device spectrum of a q x q PSF centred in a p x p frame (p defaults to q; p = 2q pads like pad_array)."""
    src, _ = as_device(psf)
    if src.ndim != 2 or src.shape[0] != src.shape[1]:
        raise ValueError("PSF must be square, got %s" % (src.shape,))
    q = src.shape[0]
    p = q if p is None else int(p)
    lib = _lib.lib()
    spec = DeviceArray((lib.pms_fft_spectrum_bytes(p),), np.uint8)
    work = Workspace.get(lib.pms_fft_work_bytes(p, p))
    check(lib.pms_psf_spectrum(src.ptr, q, p, spec.ptr, work.ptr, None))
    return PsfSpectrum(p, spec)


def gpu_r2c_fft(in1, is_gpuarray=False, store_on_gpu=False):
    """rfft2 of a square power-of-two frame (pymoresane/iuwt_convolution.py:76-106).

    With ``store_on_gpu`` the result stays in HBM as a ``PsfSpectrum`` usable by
    ``fft_convolve``; otherwise the numpy rfft2 layout is returned.
    """
    spec = psf_spectrum(in1)
    return spec if store_on_gpu else spec.to_rfft2()


def fft_convolve(in1, in2, conv_device="hip", conv_mode="linear", store_on_gpu=False):
    """Convolve in1 with the PSF whose spectrum is in2 (pymoresane/iuwt_convolution.py:18-73)."""
    check_mode(conv_device, "conv_device")
    if conv_mode not in ("linear", "circular"):
        raise ValueError("conv_mode must be 'linear' or 'circular', got %r" % (conv_mode,))
    src, _ = as_device(in1)
    if src.ndim != 2 or src.shape[0] != src.shape[1]:
        raise ValueError("fft_convolve expects a square image, got %s" % (src.shape,))
    n = src.shape[0]
    p = 2 * n if conv_mode == "linear" else n
    spec = in2 if isinstance(in2, PsfSpectrum) else PsfSpectrum.from_rfft2(in2)
    if spec.p != p:
        raise ValueError("spectrum is for a %d x %d frame but %s mode on %d x %d needs %d"
                         % (spec.p, spec.p, conv_mode, n, n, p))
    lib = _lib.lib()
    out = DeviceArray((n, n))
    work = Workspace.get(lib.pms_fft_work_bytes(n, p))
    check(lib.pms_fft_convolve(src.ptr, n, p, spec.ptr, out.ptr, work.ptr, None))
    return out if store_on_gpu else out.get()
