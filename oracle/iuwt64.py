"""Float64 restatement of the reference serial IUWT (test infrastructure only).

Follows pymoresane/iuwt.py:66-191 (``ser_iuwt_decomposition``,
``ser_iuwt_recomposition``, ``ser_a_trous``).

Semantics restated:

* The smoothing filter is the B3 spline ``[1, 4, 6, 4, 1] / 16`` (iuwt.py:82),
  applied separably, axis 0 first and then axis 1 (iuwt.py:163-189), with its
  taps dilated to ``0, +-2**s, +-2**(s+1)`` at scale ``s``.
* Out-of-range taps use the half-sample symmetric extension
  ``... c b a | a b c ... x y z | z y x ...`` (iuwt.py:166, 169, 172, 175)
  for every tap offset ``o`` with ``|o| <= side``: there the reference's
  slices reflect once.  A tap offset longer than the axis (``2**s`` or
  ``2**(s+1)`` above the side, which happens when a scale count chosen for the
  whole image is applied to a smaller subregion) makes numpy clip both slices
  of that tap to the whole axis, read in reverse: every output ``i`` then
  takes input ``side - 1 - i`` for that tap.  The restatement reproduces that
  too (``tap_index``).
* Decomposition (iuwt.py:94-106): ``scale_adjust`` pre-smoothings, then per
  scale ``C = h_i*C0; C1 = h_i*C; w_i = C0 - C1; C0 = C``.
* Recomposition (iuwt.py:131-147): ``rec = smoothed or 0``, then
  ``rec = h_i*rec + w_i`` from the coarsest kept scale down to
  ``scale_adjust``, then one plain smoothing per skipped scale.
"""

import numpy as np

B3 = np.array([1.0, 4.0, 6.0, 4.0, 1.0], dtype=np.float64) / 16.0


def reflect_index(idx, n):
    """Map any integer index onto [0, n) by the half-sample symmetric extension.

    The extension is periodic with period 2n: positions n..2n-1 mirror
    n-1..0.  Used for every out-of-range tap of the a trous filter.
    """
    m = np.mod(idx, 2 * n)
    return np.where(m >= n, 2 * n - 1 - m, m)


def tap_index(idx, off, n):
    """Input index of the tap at offset ``off`` for outputs ``idx`` on an axis of length ``n``.

    ``|off| <= n``: the half-sample symmetric extension (one reflection suffices).
    ``|off| > n``: the whole axis reversed, ``n - 1 - idx`` (numpy's clipped slices at
    iuwt.py:166-187 when ``2**s`` or ``2**(s+1)`` exceeds the side).
    """
    if abs(off) > n:
        return n - 1 - idx
    return reflect_index(idx + off, n)


def check_scale(shape, scale):
    """Raise unless every tap of scale ``scale`` reflects once (the strip-parallel restatement's range)."""
    d2 = 2 << scale
    if d2 > min(shape[-2], shape[-1]):
        raise ValueError(
            "scale %d needs 2**(scale+1)=%d <= image side %s" % (scale, d2, tuple(shape[-2:])))


def in_reflect_range(shape, scale):
    """True when every tap of scale ``scale`` is within one side (``2**(scale+1) <= side``)."""
    return (2 << scale) <= min(shape[-2], shape[-1])


_SCIPY_BACKEND = False


def use_scipy_backend(flag=True):
    """Route ``a_trous`` through ``scipy.ndimage.correlate1d`` (mode='reflect' on both axes).

    That library routine is the pin of the restatement below (equal to ~1e-16, checked in
    tests/test_oracle_pins.py) and is several times faster on large planes; long oracle runs
    (the 2048^2 by-scale case) use it.
    """
    global _SCIPY_BACKEND
    _SCIPY_BACKEND = bool(flag)


def _a_trous_scipy(c0, scale):
    from scipy import ndimage
    d = 1 << scale
    k = np.zeros(4 * d + 1)
    k[::d] = B3
    v = ndimage.correlate1d(c0, k, axis=0, mode="reflect")
    return ndimage.correlate1d(v, k, axis=1, mode="reflect")


def a_trous(c0, scale):
    """One B3 a trous smoothing at dilation 2**scale (iuwt.py:151-191)."""
    c0 = np.asarray(c0, dtype=np.float64)
    if scale < 0:
        raise ValueError("negative scale")
    if _SCIPY_BACKEND and in_reflect_range(c0.shape, scale):
        return _a_trous_scipy(c0, scale)
    d = 1 << scale
    h, w = c0.shape
    taps = [(k * d, B3[k + 2]) for k in range(-2, 3)]
    rows = np.arange(h)
    vert = np.zeros((h, w))
    for off, coeff in taps:
        vert += coeff * c0[tap_index(rows, off, h), :]
    cols = np.arange(w)
    out = np.zeros((h, w))
    for off, coeff in taps:
        out += coeff * vert[:, tap_index(cols, off, w)]
    return out


def iuwt_decomposition(in1, scale_count, scale_adjust=0, store_smoothed=False):
    """Detail planes for scales scale_adjust..scale_count-1 (iuwt.py:66-111)."""
    c0 = np.asarray(in1, dtype=np.float64)
    if scale_count <= scale_adjust:
        raise ValueError("scale_count must exceed scale_adjust")
    for s in range(scale_adjust):
        c0 = a_trous(c0, s)
    detail = np.empty((scale_count - scale_adjust,) + c0.shape)
    for s in range(scale_adjust, scale_count):
        c = a_trous(c0, s)
        detail[s - scale_adjust] = c0 - a_trous(c, s)
        c0 = c
    if store_smoothed:
        return detail, c0
    return detail


def iuwt_recomposition(in1, scale_adjust=0, smoothed_array=None):
    """Inverse of iuwt_decomposition (iuwt.py:113-149)."""
    in1 = np.asarray(in1, dtype=np.float64)
    top = in1.shape[0] + scale_adjust
    if smoothed_array is None:
        rec = np.zeros(in1.shape[1:])
    else:
        rec = np.asarray(smoothed_array, dtype=np.float64)
    for s in range(top - 1, scale_adjust - 1, -1):
        rec = a_trous(rec, s) + in1[s - scale_adjust]
    for s in range(scale_adjust - 1, -1, -1):
        rec = a_trous(rec, s)
    return rec
