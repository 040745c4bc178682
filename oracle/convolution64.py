"""Float64 restatement of the reference FFT convolution (test infrastructure only).

Follows pymoresane/iuwt_convolution.py:18-73 (cpu branch of ``fft_convolve``)
and 142-158 (``pad_array``), and the PSF-spectrum precompute of
pymoresane/main.py:175-213.

* linear:   ``fftshift(irfft2(S * rfft2(pad(x))))[N/2:3N/2, N/2:3N/2]`` where
  ``S = rfft2(psf_2N)``; closed form ``fftconvolve(x, psf_2N)[N:2N, N:2N]``.
* circular: ``fftshift(irfft2(S * rfft2(x)))`` with ``S = rfft2(psf_N)``;
  closed form ``out[i] = sum_u x[u] psf[(i - u + N/2) mod N]``.

Everything is complex128/float64 here (the reference returns float32 for the
linear branch, iuwt_convolution.py:70; the oracle keeps float64).
"""

import numpy as np

_WORKERS = 1


def set_fft_workers(n):
    """Run the 2-D transforms through scipy.fft with ``n`` threads (same pocketfft algorithm;
    results agree with numpy.fft to ~1e-16, checked in tests/test_oracle_pins.py)."""
    global _WORKERS
    _WORKERS = max(1, int(n))


def _rfft2(a):
    if _WORKERS > 1:
        import scipy.fft
        return scipy.fft.rfft2(a, workers=_WORKERS)
    return np.fft.rfft2(a)


def _irfft2(a):
    if _WORKERS > 1:
        import scipy.fft
        return scipy.fft.irfft2(a, workers=_WORKERS)
    return np.fft.irfft2(a)


def pad_array(in1):
    """Zero frame of twice the size with ``in1`` in the centre (iuwt_convolution.py:142-158)."""
    in1 = np.asarray(in1, dtype=np.float64)
    h, w = in1.shape
    out = np.zeros((2 * h, 2 * w))
    out[h // 2:h // 2 + h, w // 2:w // 2 + w] = in1
    return out


def psf_spectrum(psf_frame):
    """rfft2 of a PSF frame (main.py:178-213), complex128."""
    return _rfft2(np.asarray(psf_frame, dtype=np.float64))


def fft_convolve(in1, in2, conv_mode="linear"):
    """Convolve ``in1`` with the PSF whose rfft2 is ``in2`` (iuwt_convolution.py:62-73)."""
    in1 = np.asarray(in1, dtype=np.float64)
    if conv_mode == "linear":
        h, w = in1.shape
        full = np.fft.fftshift(_irfft2(in2 * _rfft2(pad_array(in1))))
        return full[h // 2:h // 2 + h, w // 2:w // 2 + w]
    if conv_mode == "circular":
        return np.fft.fftshift(_irfft2(in2 * _rfft2(in1)))
    raise ValueError("conv_mode must be 'linear' or 'circular', got %r" % (conv_mode,))
