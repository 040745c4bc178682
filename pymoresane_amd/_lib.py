"""ctypes binding of libpmship.so (include/pymoresane_hip.h) and device-array plumbing.

The library is the only compute path of this package: if it is missing or no
GPU is visible, every operation raises ``PmsError``.  There is no CPU
fallback.
"""

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PMS_LIBRARY", os.path.join(_HERE, "libpmship.so"))

c_int, c_size_t, c_double, c_float, c_void_p = ctypes.c_int, ctypes.c_size_t, ctypes.c_double, ctypes.c_float, ctypes.c_void_p
c_int_p = ctypes.POINTER(ctypes.c_int)
c_float_p = ctypes.POINTER(ctypes.c_float)


class PmsError(RuntimeError):
    """A libpmship.so call failed; the message is the library's pms_last_error()."""

    def __init__(self, code, msg):
        super().__init__("%s (status %d)" % (msg, code))
        self.code = code


class TraceEvent(ctypes.Structure):
    _fields_ = [("kind", c_int), ("major", c_int), ("a", c_int), ("b", c_int), ("c", c_int), ("pad", c_int),
                ("x", c_double), ("y", c_double)]


class MoresaneParams(ctypes.Structure):
    _fields_ = [("n", c_int), ("psf_n", c_int), ("subregion", c_int), ("scale_count", c_int),
                ("sigma_level", c_double), ("loop_gain", c_double), ("tolerance", c_double), ("accuracy", c_double),
                ("major_loop_miter", c_int), ("minor_loop_miter", c_int), ("conv_mode", c_int),
                ("enforce_positivity", c_int), ("edge_suppression", c_int), ("edge_offset", c_int),
                ("flux_threshold", c_double), ("neg_comp", c_int), ("edge_excl", c_int), ("int_excl", c_int),
                ("precision", c_int)]


class BatchJob(ctypes.Structure):
    _fields_ = [("dirty", c_void_p), ("psf", c_void_p), ("mask", c_void_p), ("model", c_void_p),
                ("residual", c_void_p), ("prm", MoresaneParams), ("by_scale", c_int), ("start_scale", c_int),
                ("stop_scale", c_int), ("status", c_int), ("major_iters", c_int), ("complete", c_int),
                ("seconds", c_double), ("t_start", c_double), ("wait_seconds", c_double)]


class ServeRequest(ctypes.Structure):
    _fields_ = [("dirty_host", c_void_p), ("psf_host", c_void_p), ("mask_host", c_void_p), ("model_host", c_void_p),
                ("residual_host", c_void_p), ("prm", MoresaneParams), ("by_scale", c_int), ("start_scale", c_int),
                ("stop_scale", c_int), ("status", c_int), ("major_iters", c_int), ("pad", c_int),
                ("footprint", c_size_t), ("t_submit", c_double), ("t_start", c_double), ("t_done", c_double)]


EV_SCALES, EV_MINOR, EV_UPDATE, EV_NOCOEFF, EV_POOR, EV_NOWORK, EV_BYSCALE = 1, 2, 3, 4, 5, 6, 7

# name -> (restype, argtypes); mirrors include/pymoresane_hip.h one to one
SIGNATURES = {
    "pms_init": (c_int, [c_int]),
    "pms_last_error": (ctypes.c_char_p, []),
    "pms_get_device_count": (c_int, [c_int_p]),
    "pms_device_synchronize": (c_int, []),
    "pms_malloc": (c_int, [ctypes.POINTER(c_void_p), c_size_t]),
    "pms_free": (c_int, [c_void_p]),
    "pms_memcpy_h2d": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p]),
    "pms_memcpy_d2h": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p]),
    "pms_memcpy_d2d": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p]),
    "pms_fill_f32": (c_int, [c_void_p, c_size_t, c_float, c_void_p]),
    "pms_copy_f32": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p]),
    "pms_stream_create": (c_int, [ctypes.POINTER(c_void_p)]),
    "pms_stream_destroy": (c_int, [c_void_p]),
    "pms_stream_synchronize": (c_int, [c_void_p]),
    "pms_event_create": (c_int, [ctypes.POINTER(c_void_p)]),
    "pms_event_destroy": (c_int, [c_void_p]),
    "pms_event_record": (c_int, [c_void_p, c_void_p]),
    "pms_event_elapsed_ms": (c_int, [c_void_p, c_void_p, c_float_p]),
    "pms_iuwt_work_bytes": (c_size_t, [c_int, c_int]),
    "pms_iuwt_decompose": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "pms_iuwt_recompose": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "pms_fft_spectrum_bytes": (c_size_t, [c_int]),
    "pms_fft_work_bytes": (c_size_t, [c_int, c_int]),
    "pms_psf_spectrum": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "pms_spectrum_from_rfft2": (c_int, [c_void_p, c_int, c_void_p]),
    "pms_fft_convolve": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "pms_threshold_work_bytes": (c_size_t, [c_int, c_int, c_int]),
    "pms_estimate_threshold": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "pms_estimate_threshold_noalloc": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p]),
    "pms_apply_threshold": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_double, c_void_p, c_void_p, c_void_p]),
    "pms_extract_work_bytes": (c_size_t, [c_int, c_int, c_int]),
    "pms_source_extraction": (c_int, [c_void_p, c_int, c_int, c_int, c_double, c_int, c_void_p, c_void_p, c_void_p,
                                      c_void_p, c_void_p]),
    "pms_snr_work_bytes": (c_size_t, [c_size_t]),
    "pms_snr_ratio": (c_int, [c_void_p, c_void_p, c_size_t, ctypes.POINTER(c_double), c_void_p, c_void_p]),
    "pms_moresane": (c_int, [ctypes.POINTER(MoresaneParams), c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                             c_int_p, ctypes.POINTER(TraceEvent), c_int, c_int_p, c_void_p]),
    "pms_moresane_by_scale": (c_int, [ctypes.POINTER(MoresaneParams), c_int, c_int, c_void_p, c_void_p, c_void_p,
                                      c_void_p, c_void_p, ctypes.POINTER(TraceEvent), c_int, c_int_p, c_void_p]),
    "pms_batch_run": (c_int, [ctypes.POINTER(BatchJob), c_int, c_int, c_int]),
    "pms_server_create": (c_int, [c_int, c_double, ctypes.POINTER(c_void_p)]),
    "pms_server_submit": (c_int, [c_void_p, ctypes.POINTER(ServeRequest)]),
    "pms_server_wait": (c_int, [c_void_p, ctypes.POINTER(ServeRequest), c_double, c_int_p]),
    "pms_server_drain": (c_int, [c_void_p]),
    "pms_server_now": (c_int, [c_void_p, ctypes.POINTER(c_double)]),
    "pms_server_destroy": (c_int, [c_void_p]),
}

_lock = threading.Lock()
_lib = None
_initialised = False


def load():
    """dlopen libpmship.so and bind every C-ABI symbol (no GPU needed)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise PmsError(-4, "HIP library %s is not built (run __graft_entry__.build())" % LIB_PATH)
        lib_ = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib_, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib_
        return _lib


def device_count():
    n = c_int(0)
    check(load().pms_get_device_count(ctypes.byref(n)))
    return n.value


def gpu_available():
    try:
        return device_count() > 0
    except (PmsError, OSError):
        return False


def check(status):
    if status != 0:
        msg = _lib.pms_last_error() if _lib is not None else b"library not loaded"
        raise PmsError(status, msg.decode("utf-8", "replace") if msg else "unknown error")
    return status


def lib():
    """The bound library, initialised on device 0; raises PmsError without a GPU."""
    global _initialised
    lib_ = load()
    if not _initialised:
        with _lock:
            if not _initialised:
                n = c_int(0)
                lib_.pms_get_device_count(ctypes.byref(n))
                if n.value <= 0:
                    raise PmsError(-4, "no HIP device visible: the MORESANE HIP path needs an MI355X (gfx950)")
                check(lib_.pms_init(0))
                _initialised = True
    return lib_


# ---------------------------------------------------------------------------------------- device arrays
class DeviceArray(object):
    """A device buffer with a numpy-like shape/dtype; ``get()`` copies it to the host (like pycuda's gpuarray)."""

    def __init__(self, shape, dtype=np.float32):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if not isinstance(shape, tuple) else tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self._ptr = c_void_p(0)
        self._owner = True
        if self.nbytes:
            check(lib().pms_malloc(ctypes.byref(self._ptr), self.nbytes))

    @classmethod
    def view(cls, base, offset_bytes, shape, dtype):
        """Non-owning view into another DeviceArray's memory."""
        obj = cls.__new__(cls)
        obj.shape = tuple(shape)
        obj.dtype = np.dtype(dtype)
        obj.nbytes = int(np.prod(obj.shape, dtype=np.int64)) * obj.dtype.itemsize
        if offset_bytes + obj.nbytes > base.nbytes:
            raise ValueError("view out of range")
        obj._ptr = c_void_p((base.ptr or 0) + offset_bytes)
        obj._owner = False
        obj._base = base
        return obj

    @property
    def ptr(self):
        return self._ptr.value

    @property
    def size(self):
        return int(np.prod(self.shape, dtype=np.int64))

    @property
    def ndim(self):
        return len(self.shape)

    @classmethod
    def from_host(cls, arr, dtype=None):
        arr = np.ascontiguousarray(arr if dtype is None else np.asarray(arr, dtype=dtype))
        out = cls(arr.shape, arr.dtype)
        out.set(arr)
        return out

    def set(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        if arr.nbytes != self.nbytes:
            raise ValueError("size mismatch: %d vs %d bytes" % (arr.nbytes, self.nbytes))
        if self.nbytes:
            check(lib().pms_memcpy_h2d(self.ptr, arr.ctypes.data, self.nbytes, None))

    def get(self):
        out = np.empty(self.shape, dtype=self.dtype)
        if self.nbytes:
            check(lib().pms_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes, None))
        return out

    def fill(self, value):
        if self.dtype != np.float32:
            raise TypeError("fill supports float32 buffers")
        check(lib().pms_fill_f32(self.ptr, self.size, float(value), None))

    def free(self):
        if self._owner and self._ptr.value:
            try:
                check(_lib.pms_free(self._ptr))
            except Exception:
                pass
        self._ptr = c_void_p(0)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __repr__(self):
        return "DeviceArray(shape=%s, dtype=%s)" % (self.shape, self.dtype)


class Workspace(object):
    """Grow-only scratch buffer reused across calls (one per thread)."""

    _tls = threading.local()

    @classmethod
    def get(cls, nbytes):
        buf = getattr(cls._tls, "buf", None)
        if buf is None or buf.nbytes < nbytes:
            if buf is not None:
                buf.free()
            buf = DeviceArray((max(int(nbytes), 256),), np.uint8)
            cls._tls.buf = buf
        return buf


def as_device(a, dtype=np.float32):
    """(DeviceArray, uploaded?) for a numpy array or an existing DeviceArray."""
    if isinstance(a, DeviceArray):
        if a.dtype != np.dtype(dtype):
            raise TypeError("device array has dtype %s, expected %s" % (a.dtype, np.dtype(dtype)))
        return a, False
    return DeviceArray.from_host(np.asarray(a, dtype=dtype)), True


class Event(object):
    def __init__(self):
        self._e = c_void_p(0)
        check(lib().pms_event_create(ctypes.byref(self._e)))

    def record(self, stream=None):
        check(lib().pms_event_record(self._e, stream))

    def elapsed_ms(self, later):
        ms = c_float(0.0)
        check(lib().pms_event_elapsed_ms(self._e, later._e, ctypes.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            if self._e.value:
                _lib.pms_event_destroy(self._e)
        except Exception:
            pass


def synchronize():
    check(lib().pms_device_synchronize())
