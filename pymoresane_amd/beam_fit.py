"""Restoring-beam fit (pymoresane/beam_fit.py:6-69), run on the host.

The beam is an elliptical Gaussian fitted by least squares to the main lobe
of the peak-normalised PSF: the 8-connected region of values above 0.5 that
contains the PSF maximum.  It runs once per restore, off the deconvolution
loop, so it stays on the CPU (scipy); the restoring convolution itself runs on
the GPU.  The reference's crop test ``psf.shape > 512`` compares a tuple with
an int (always true under Python 2); here the central 512 x 512 crop is taken
only when the PSF is larger than 512.
"""

import numpy as np
from scipy import ndimage
from scipy.optimize import curve_fit


def _gauss(xy, amp, sx, sy, theta):
    u = xy[:, 0] * np.cos(theta) - xy[:, 1] * np.sin(theta)
    v = xy[:, 0] * np.sin(theta) + xy[:, 1] * np.cos(theta)
    return amp * np.exp(-(u ** 2 / (2.0 * sx ** 2) + v ** 2 / (2.0 * sy ** 2)))


def beam_fit(psf, cdelt1, cdelt2):
    """(clean_beam like psf, [bmaj, bmin, bpa]) with widths in the units of cdelt."""
    psf = np.asarray(psf, dtype=np.float64)
    h, w = psf.shape
    if max(h, w) > 512:
        r0, c0 = h // 2 - 256, w // 2 - 256
        win = (slice(r0, r0 + 512), slice(c0, c0 + 512))
    else:
        win = (slice(0, h), slice(0, w))
    centre = psf[win] / np.max(psf[win])
    peak = np.unravel_index(np.argmax(centre), centre.shape)
    labels, _ = ndimage.label(centre > 0.5)
    lobe = np.where(labels == labels[peak], centre, 0.0)
    xs = np.arange(centre.shape[1]) - peak[1]
    ys = -(np.arange(centre.shape[0]) - peak[0])
    gx, gy = np.meshgrid(xs, ys)
    xy = np.column_stack((gx.ravel(), gy.ravel()))
    opt = curve_fit(_gauss, xy, lobe.ravel(), p0=(1.0, 1.0, 1.0, 0.0))[0]
    beam = np.zeros_like(psf)
    beam[win] = _gauss(xy, *opt).reshape(centre.shape)
    beam /= np.max(beam)
    fwhm = 2.0 * np.sqrt(2.0 * np.log(2.0))
    bmaj = fwhm * max(opt[1], opt[2]) * cdelt1
    bmin = fwhm * min(opt[1], opt[2]) * cdelt2
    bpa = np.degrees(opt[3]) % 360 - 90
    return beam, [abs(bmaj), abs(bmin), bpa]
