"""Float64 restatement of the reference toolbox (test infrastructure only).

Follows pymoresane/iuwt_toolbox.py:17-154 and 292-306.

* ``estimate_threshold`` (17-48): per scale ``median(|w[mask]|) / 0.6745``;
  the optional mask drops an ``edge_excl`` border and a square of
  half-width ``int_excl`` around ``mid = in1.shape[1] // 2`` -- the row count
  of the (scales, rows, cols) stack, used for both axes (29 and 40), so on a
  non-square plane the square sits on the diagonal, not at the centre.
* ``apply_threshold`` (50-75): keeps ``w`` where ``|w| > sigma_level * t``
  (strict), negatives included.
* ``source_extraction`` (97-154): 8-connected labels of the non-zero pixels
  of every scale (negative values are foreground); from the coarsest scale
  down, an object is kept when one of its pixels has ``w >= tol * max(w)``
  (``|w| >= tol * max|w|`` with ``neg_comp``) and, below the coarsest
  scale, lies inside the kept mask of the next coarser scale.  Returns
  ``(mask * w, mask)``.
* ``snr_ratio`` (292-306): ``20 log10(||a|| / ||a - b||)``.
"""

import numpy as np
from scipy import ndimage

EIGHT_CONNECTED = np.ones((3, 3), dtype=int)


def noise_mask(shape, edge_excl=0, int_excl=0):
    """Boolean mask of the pixels that enter the MAD estimate (iuwt_toolbox.py:31-43)."""
    h, w = shape
    mid = h // 2  # in1.shape[1] of the (scales, rows, cols) stack, for both axes
    mask = np.ones((h, w), dtype=bool)
    if edge_excl != 0:
        mask[:] = False
        mask[edge_excl:-edge_excl, edge_excl:-edge_excl] = True
    if int_excl != 0:
        mask[mid - int_excl:mid + int_excl, mid - int_excl:mid + int_excl] = False
    return mask


def estimate_threshold(in1, edge_excl=0, int_excl=0):
    in1 = np.asarray(in1, dtype=np.float64)
    sel = noise_mask(in1.shape[1:], edge_excl, int_excl)
    return np.array([np.median(np.abs(in1[i][sel])) / 0.6745 for i in range(in1.shape[0])])


def apply_threshold(in1, threshold, sigma_level=4):
    in1 = np.asarray(in1, dtype=np.float64)
    if in1.ndim == 2:
        return np.where(np.abs(in1) > sigma_level * threshold, in1, 0.0)
    out = np.empty_like(in1)
    for i in range(in1.shape[0]):
        out[i] = np.where(np.abs(in1[i]) > sigma_level * threshold[i], in1[i], 0.0)
    return out


def label_scale(plane):
    """8-connected component labels of the non-zero pixels, raster order of first pixel."""
    labels, count = ndimage.label(plane != 0, structure=EIGHT_CONNECTED)
    return labels.astype(np.int32), int(count)


def source_extraction(in1, tolerance, neg_comp=False, return_labels=False):
    in1 = np.asarray(in1, dtype=np.float64)
    ns = in1.shape[0]
    labels = np.empty(in1.shape, dtype=np.int32)
    maxima = np.empty(ns)
    for i in range(ns):
        maxima[i] = np.max(np.abs(in1[i])) if neg_comp else np.max(in1[i])
        labels[i], _ = label_scale(in1[i])
    mask = np.zeros(in1.shape, dtype=np.int32)
    for i in range(ns - 1, -1, -1):
        vals = np.abs(in1[i]) if neg_comp else in1[i]
        seed = (vals >= tolerance * maxima[i]) & (labels[i] > 0)
        if i < ns - 1:
            seed &= mask[i + 1] != 0
        kept = np.unique(labels[i][seed])
        mask[i] = np.isin(labels[i], kept[kept > 0]).astype(np.int32)
    if return_labels:
        return mask * in1, mask, labels
    return mask * in1, mask


def snr_ratio(in1, in2):
    in1 = np.asarray(in1, dtype=np.float64)
    in2 = np.asarray(in2, dtype=np.float64)
    return 20.0 * np.log10(np.linalg.norm(in1) / np.linalg.norm(in1 - in2))
