"""Float64 restatement of the reference MORESANE driver (test infrastructure only).

Follows pymoresane/main.py:18-49 (``DataImage``), 53-521 (``moresane``) and
523-599 (``moresane_by_scale``), with every array in float64 (the reference
keeps dirty/PSF/model/residual in float32, main.py:34-35, 238).  Integer
division replaces the Python 2 ``/`` in slice bounds.  Per-scale PSF
energies and normalised maxima are kept in float64 (the reference stores
them in float32 arrays, main.py:223, 297).

The driver records a decision trace -- a list of tuples -- that the HIP
engine reproduces event for event:

* ``("scales", major_iter, min_scale, max_scale, scale_adjust, max_coeff)``
* ``("minor", major_iter, n_iter, exit_code, snr)`` with exit codes
  0 = iteration cap, 1 = first-iteration SNR above 40 (min_scale += 1),
  2 = SNR above 40 (accept), 3 = SNR fell, above 10.5 (keep previous x),
  4 = SNR fell, too small (min_scale += 1)
* ``("update", major_iter, std_current, std_ratio, reverted)``
* ``("nocoeff", major_iter)`` / ``("poor", major_iter)`` / ``("nowork",)``
"""

import numpy as np

from . import convolution64 as conv
from . import iuwt64 as iuwt
from . import toolbox64 as tools

EXIT_CAP, EXIT_FIRST_HIGH, EXIT_ACCEPT, EXIT_FELL_OK, EXIT_FELL_LOW = 0, 1, 2, 3, 4


def psf_spectra(psf, dirty_n, subregion, conv_mode):
    """PSF spectra for the minor loop and the residual (main.py:175-213, cpu branch)."""
    psf = np.asarray(psf, dtype=np.float64)
    pn = psf.shape[0]
    double = pn == 2 * dirty_n
    full = subregion == dirty_n
    if conv_mode == "circular":
        if double:
            lo = pn // 2 - subregion // 2
            sub = conv.psf_spectrum(psf[lo:lo + subregion, lo:lo + subregion])
            lo = pn // 2 - dirty_n // 2
            return sub, conv.psf_spectrum(psf[lo:lo + dirty_n, lo:lo + dirty_n])
        lo = dirty_n // 2 - subregion // 2
        psf_sub = psf[lo:lo + subregion, lo:lo + subregion]
        sub = conv.psf_spectrum(psf_sub)
        if psf_sub.shape == psf.shape:
            return sub, sub
        return sub, conv.psf_spectrum(psf)
    if conv_mode == "linear":
        if double:
            if full:
                s = conv.psf_spectrum(psf)
                return s, s
            lo = pn // 2 - subregion
            return (conv.psf_spectrum(psf[lo:lo + 2 * subregion, lo:lo + 2 * subregion]),
                    conv.psf_spectrum(psf))
        if full:
            s = conv.psf_spectrum(conv.pad_array(psf))
            return s, s
        lo = pn // 2 - subregion
        return (conv.psf_spectrum(psf[lo:lo + 2 * subregion, lo:lo + 2 * subregion]),
                conv.psf_spectrum(conv.pad_array(psf)))
    raise ValueError("conv_mode must be 'linear' or 'circular'")


def suppression_borders(scale_count, edge_suppression, edge_offset):
    """Per-scale border width of the edge-suppression mask (main.py:249-260), or None."""
    if edge_suppression:
        borders, corruption = [], 0
        for i in range(scale_count):
            corruption += 2 * 2 ** i
            borders.append(edge_offset if edge_offset > corruption else corruption)
        return borders
    if edge_offset > 0:
        return [edge_offset] * scale_count
    return None


def border_mask(n, b):
    m = np.zeros((n, n))
    m[b:-b, b:-b] = 1.0
    return m


class OracleImage(object):
    """float64 counterpart of ``DataImage`` (main.py:18-49)."""

    def __init__(self, dirty_data, psf_data, mask_data=None):
        self.dirty_data = np.asarray(dirty_data, dtype=np.float64).copy()
        self.psf_data = np.asarray(psf_data, dtype=np.float64).copy()
        self.mask = None if mask_data is None else np.asarray(mask_data, dtype=np.float64)
        self.dirty_data_shape = self.dirty_data.shape
        self.psf_data_shape = self.psf_data.shape
        self.complete = False
        self.model = np.zeros_like(self.dirty_data)
        self.residual = self.dirty_data.copy()
        self.trace = []

    def moresane(self, subregion=None, scale_count=None, sigma_level=4, loop_gain=0.1, tolerance=0.75,
                 accuracy=1e-6, major_loop_miter=100, minor_loop_miter=30, conv_mode="linear",
                 enforce_positivity=False, edge_suppression=False, edge_offset=0, flux_threshold=0,
                 neg_comp=False, edge_excl=0, int_excl=0):
        n = self.dirty_data_shape[0]
        if n % 2 == 1:
            raise ValueError("Image size is uneven. Please use even dimensions.")
        if subregion is None or subregion > n:
            subregion = n
        if scale_count is None or scale_count > (np.log2(n) - 1):
            scale_count = int(np.log2(n) - 1)
        trace = self.trace

        lo = n // 2 - subregion // 2
        sub = (slice(lo, lo + subregion), slice(lo, lo + subregion))
        dirty_sub = self.dirty_data[sub]
        pn = self.psf_data_shape[0]
        if pn == 2 * n:
            plo = pn // 2 - subregion // 2
            psf_sub = self.psf_data[plo:plo + subregion, plo:plo + subregion]
        else:
            psf_sub = self.psf_data[sub]
        psf_sub_fft, psf_data_fft = psf_spectra(self.psf_data, n, subregion, conv_mode)

        psf_dec = iuwt.iuwt_decomposition(psf_sub, scale_count)
        energies = np.sqrt(np.sum(psf_dec ** 2, axis=(1, 2)))

        major = 0
        max_coeff = 1.0
        model = np.zeros_like(self.dirty_data)
        residual = None
        std_current, std_ratio = 1000.0, 1.0
        min_scale = 0
        borders = suppression_borders(scale_count, edge_suppression, edge_offset)
        nsm = None
        dirty_thresh = None
        x = None

        def op(v, max_scale, scale_adjust, emask):
            return emask * iuwt.iuwt_decomposition(conv.fft_convolve(v, psf_sub_fft, conv_mode),
                                                   max_scale, scale_adjust)

        while (major < major_loop_miter and max_coeff > 0 and std_ratio > accuracy
               and np.max(dirty_sub) > flux_threshold):
            while min_scale < scale_count:
                if min_scale == 0:
                    dec = iuwt.iuwt_decomposition(dirty_sub, scale_count)
                    thresholds = tools.estimate_threshold(dec, edge_excl, int_excl)
                    if self.mask is not None:
                        dec = iuwt.iuwt_decomposition(dirty_sub * self.mask[sub], scale_count)
                    dirty_thresh = tools.apply_threshold(dec, thresholds, sigma_level)
                    if borders is not None:
                        for i in range(scale_count):
                            dirty_thresh[i] *= border_mask(subregion, borders[i])
                    nsm = np.array([np.max(dirty_thresh[i]) / energies[i] for i in range(scale_count)])

                max_index = int(np.argmax(nsm[min_scale:])) + min_scale
                max_scale = max_index + 1
                max_coeff = float(nsm[max_index])
                if max_coeff == 0:
                    trace.append(("nocoeff", major))
                    break
                scale_adjust = 0
                for i in range(max_index - 1, -1, -1):
                    if nsm[i] == 0:
                        scale_adjust = i + 1
                        break
                trace.append(("scales", major, min_scale, max_scale, scale_adjust, max_coeff))

                extracted, emask = tools.source_extraction(dirty_thresh[scale_adjust:max_scale], tolerance,
                                                           neg_comp=neg_comp)
                recomposed = iuwt.iuwt_recomposition(extracted, scale_adjust)

                x = np.zeros_like(recomposed)
                r = recomposed.copy()
                p = recomposed.copy()
                niter = 0
                snr_last = snr_current = 0.0
                exit_code = EXIT_CAP
                while niter < minor_loop_miter:
                    ap = iuwt.iuwt_recomposition(op(p, max_scale, scale_adjust, emask), scale_adjust)
                    alpha = np.sum(r * r) / np.sum(p * ap)
                    xn = x + alpha * p
                    if enforce_positivity and np.min(xn) < 0:
                        xn[xn < 0] = 0
                        p = (xn - x) / alpha
                        ap = iuwt.iuwt_recomposition(op(p, max_scale, scale_adjust, emask), scale_adjust)
                    rn = r - alpha * ap
                    beta = np.sum(rn * rn) / np.sum(r * r)
                    p = rn + beta * p
                    model_sources = op(xn, max_scale, scale_adjust, emask)
                    snr_last = snr_current
                    snr_current = tools.snr_ratio(extracted, model_sources)
                    niter += 1
                    if niter == 1 and snr_current > 40:
                        min_scale += 1
                        exit_code = EXIT_FIRST_HIGH
                        break
                    if snr_current > 40:
                        x = xn
                        min_scale = 0
                        exit_code = EXIT_ACCEPT
                        break
                    if niter > 2 and snr_current <= snr_last:
                        if snr_current > 10.5:
                            min_scale = 0
                            exit_code = EXIT_FELL_OK
                        else:
                            min_scale += 1
                            exit_code = EXIT_FELL_LOW
                        break
                    r = rn
                    x = xn
                trace.append(("minor", major, niter, exit_code, float(snr_current)))
                if niter == minor_loop_miter and snr_current > 10.5:
                    min_scale = 0
                    break
                if min_scale == 0:
                    break
                if exit_code == EXIT_CAP:
                    # main.py:274-462 would repeat this minor loop unchanged forever: min_scale > 0
                    # skips the decomposition and nothing else moves
                    raise RuntimeError("minor loop hit its cap at SNR <= 10.5 with min_scale > 0: "
                                       "the reference loops forever here")

            if min_scale == scale_count:
                trace.append(("poor", major))
                break

            if max_coeff > 0:
                model[sub] += loop_gain * x
                residual = self.dirty_data - conv.fft_convolve(model, psf_data_fft, conv_mode)
                std_last = std_current
                std_current = float(np.std(residual[sub]))
                std_ratio = (std_last - std_current) / std_last
                reverted = std_ratio < 0
                if reverted:
                    model[sub] -= loop_gain * x
                    residual = self.dirty_data - conv.fft_convolve(model, psf_data_fft, conv_mode)
                trace.append(("update", major, std_current, std_ratio, bool(reverted)))
                dirty_sub = residual[sub]
                major += 1

            if major == 0:
                trace.append(("nowork",))
                self.complete = True
                break

        if major > 0:
            self.model += model
            self.residual = residual
        return major

    def moresane_by_scale(self, start_scale=1, stop_scale=20, **kwargs):
        """main.py:523-599: repeated ``moresane`` with a growing scale count."""
        dirty = self.dirty_data
        scale_count = start_scale
        while not self.complete:
            self.trace.append(("byscale", scale_count))
            self.moresane(scale_count=scale_count, **kwargs)
            self.dirty_data = self.residual
            scale_count += 1
            if scale_count > np.log2(self.dirty_data.shape[0]) - 1:
                break
            if scale_count > stop_scale:
                break
        self.dirty_data = dirty
        self.complete = False
