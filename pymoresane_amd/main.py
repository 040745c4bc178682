"""DataImage / FitsImage with the MORESANE deconvolution running on the MI355X.

Drop-in for pymoresane/main.py:18-616: ``DataImage(dirty, psf, mask, cdelt1,
cdelt2)`` and ``FitsImage(dirty.fits, psf.fits, mask=None)`` with
``moresane(...)`` and ``moresane_by_scale(...)`` taking the reference's
keyword arguments (main.py:53-57, 523-527).  The HIP engine in libpmship.so
runs the whole major/minor loop with every array resident in HBM; this
module uploads the inputs, calls it, and reads back model, residual and the
decision trace.  ``decom_mode``, ``conv_device`` and ``extraction_mode`` must
be ``'hip'`` (``'gpu'`` is an alias; ``all_on_gpu=True`` selects it too);
``core_count`` is accepted and ignored.
"""

import ctypes
import logging

import numpy as np

from . import _lib
from . import fits as _fits
from ._lib import DeviceArray, MoresaneParams, TraceEvent, check
from .iuwt import check_mode

logger = logging.getLogger(__name__)

TRACE_CAP = 1 << 16


def decode_trace(events, n):
    """Decision trace as tuples in the oracle's format."""
    out = []
    for e in events[:n]:
        k = e.kind
        if k == _lib.EV_SCALES:
            out.append(("scales", e.major, e.a, e.b, e.c, e.x))
        elif k == _lib.EV_MINOR:
            out.append(("minor", e.major, e.a, e.b, e.x))
        elif k == _lib.EV_UPDATE:
            out.append(("update", e.major, e.x, e.y, bool(e.a)))
        elif k == _lib.EV_NOCOEFF:
            out.append(("nocoeff", e.major))
        elif k == _lib.EV_POOR:
            out.append(("poor", e.major))
        elif k == _lib.EV_NOWORK:
            out.append(("nowork",))
        elif k == _lib.EV_BYSCALE:
            out.append(("byscale", e.a))
    return out


PRECISIONS = {"float32": 0, "float64": 1}


def make_params(n, psf_n, subregion, scale_count, sigma_level, loop_gain, tolerance, accuracy, major_loop_miter,
                minor_loop_miter, conv_mode, enforce_positivity, edge_suppression, edge_offset, flux_threshold,
                neg_comp, edge_excl, int_excl, precision="float32"):
    if conv_mode not in ("linear", "circular"):
        raise ValueError("conv_mode must be 'linear' or 'circular', got %r" % (conv_mode,))
    if precision not in PRECISIONS:
        raise ValueError("precision must be 'float32' or 'float64', got %r" % (precision,))
    prm = MoresaneParams()
    prm.n, prm.psf_n = int(n), int(psf_n)
    prm.subregion = 0 if subregion is None else int(subregion)
    prm.scale_count = 0 if scale_count is None else int(scale_count)
    prm.sigma_level, prm.loop_gain = float(sigma_level), float(loop_gain)
    prm.tolerance, prm.accuracy = float(tolerance), float(accuracy)
    prm.major_loop_miter, prm.minor_loop_miter = int(major_loop_miter), int(minor_loop_miter)
    prm.conv_mode = 0 if conv_mode == "linear" else 1
    prm.enforce_positivity = int(bool(enforce_positivity))
    prm.edge_suppression = int(bool(edge_suppression))
    prm.edge_offset = int(edge_offset)
    prm.flux_threshold = float(flux_threshold)
    prm.neg_comp = int(bool(neg_comp))
    prm.edge_excl, prm.int_excl = int(edge_excl), int(int_excl)
    prm.precision = PRECISIONS[precision]
    return prm


class DataImage(object):
    """Base class managing one deconvolution (pymoresane/main.py:18-49)."""

    def __init__(self, dirty_data, psf_data, mask_data, cdelt1, cdelt2):
        self.dirty_data = np.asarray(dirty_data).astype(np.float32)
        self.psf_data = np.asarray(psf_data).astype(np.float32)
        self.mask_name = None
        self.mask = None if mask_data is None else np.asarray(mask_data, dtype=np.float32)
        self.dirty_data_shape = self.dirty_data.shape
        self.psf_data_shape = self.psf_data.shape
        self.complete = False
        self.model = np.zeros_like(self.dirty_data)
        self.residual = np.copy(self.dirty_data)
        self.restored = np.zeros_like(self.dirty_data)
        self.cdelt1 = cdelt1
        self.cdelt2 = cdelt2
        self.trace = []

    # ----------------------------------------------------------------------------- helpers
    def _check_modes(self, all_on_gpu, decom_mode, conv_device, extraction_mode):
        if all_on_gpu:
            decom_mode = conv_device = extraction_mode = "hip"
        check_mode(decom_mode, "decom_mode")
        check_mode(conv_device, "conv_device")
        check_mode(extraction_mode, "extraction_mode")

    def _check_shapes(self):
        n = self.dirty_data_shape[0]
        if len(self.dirty_data_shape) != 2 or self.dirty_data_shape[1] != n:
            raise ValueError("the dirty image must be square, got %s" % (self.dirty_data_shape,))
        if n % 2 == 1:
            logger.error("Image size is uneven. Please use even dimensions.")
            raise ValueError("Image size is uneven. Please use even dimensions.")
        pn = self.psf_data_shape[0]
        if self.psf_data_shape != (pn, pn) or pn not in (n, 2 * n):
            raise ValueError("the PSF must be square and the dirty size or twice it, got %s" % (self.psf_data_shape,))

    def _run(self, by_scale, start_scale, stop_scale, prm):
        lib = _lib.lib()
        dirty = DeviceArray.from_host(self.dirty_data)
        psf = DeviceArray.from_host(self.psf_data)
        mask = DeviceArray.from_host(self.mask) if self.mask is not None else None
        model = DeviceArray.from_host(self.model)
        resid = DeviceArray.from_host(self.residual)
        trace = (TraceEvent * TRACE_CAP)()
        tlen = ctypes.c_int(0)
        complete = ctypes.c_int(0)
        try:
            if by_scale:
                check(lib.pms_moresane_by_scale(ctypes.byref(prm), int(start_scale), int(stop_scale), dirty.ptr,
                                                psf.ptr, mask.ptr if mask else None, model.ptr, resid.ptr, trace,
                                                TRACE_CAP, ctypes.byref(tlen), None))
            else:
                check(lib.pms_moresane(ctypes.byref(prm), dirty.ptr, psf.ptr, mask.ptr if mask else None, model.ptr,
                                       resid.ptr, ctypes.byref(complete), trace, TRACE_CAP, ctypes.byref(tlen), None))
        finally:
            self.trace.extend(decode_trace(trace, min(tlen.value, TRACE_CAP)))
        self.model = model.get()
        self.residual = resid.get()
        if not by_scale:
            self.complete = bool(complete.value)
        return tlen.value

    # ----------------------------------------------------------------------------- API
    def moresane(self, subregion=None, scale_count=None, sigma_level=4, loop_gain=0.1, tolerance=0.75, accuracy=1e-6,
                 major_loop_miter=100, minor_loop_miter=30, all_on_gpu=False, decom_mode="hip", core_count=1,
                 conv_device="hip", conv_mode="linear", extraction_mode="hip", enforce_positivity=False,
                 edge_suppression=False, edge_offset=0, flux_threshold=0, neg_comp=False, edge_excl=0, int_excl=0,
                 precision="float32"):
        """Primary deconvolution (pymoresane/main.py:53-521); updates self.model and self.residual.

        ``precision`` (not in the reference): ``'float32'`` runs the wavelet planes, convolutions, CG
        vectors, model and residual in float32 with float64 inner products, step sizes and norms;
        ``'float64'`` runs all device arithmetic in float64, the arithmetic of the reference's serial
        numpy path, at about twice the memory traffic.
        """
        logger.info("Starting...")
        self._check_modes(all_on_gpu, decom_mode, conv_device, extraction_mode)
        self._check_shapes()
        prm = make_params(self.dirty_data_shape[0], self.psf_data_shape[0], subregion, scale_count, sigma_level,
                          loop_gain, tolerance, accuracy, major_loop_miter, minor_loop_miter, conv_mode,
                          enforce_positivity, edge_suppression, edge_offset, flux_threshold, neg_comp, edge_excl,
                          int_excl, precision)
        self._run(False, 0, 0, prm)

    def moresane_by_scale(self, start_scale=1, stop_scale=20, subregion=None, sigma_level=4, loop_gain=0.1,
                          tolerance=0.75, accuracy=1e-6, major_loop_miter=100, minor_loop_miter=30, all_on_gpu=False,
                          decom_mode="hip", core_count=1, conv_device="hip", conv_mode="linear",
                          extraction_mode="hip", enforce_positivity=False, edge_suppression=False, edge_offset=0,
                          flux_threshold=0, neg_comp=False, edge_excl=0, int_excl=0, precision="float32"):
        """Scale-by-scale deconvolution (pymoresane/main.py:523-599); ``precision`` as for moresane."""
        self._check_modes(all_on_gpu, decom_mode, conv_device, extraction_mode)
        self._check_shapes()
        prm = make_params(self.dirty_data_shape[0], self.psf_data_shape[0], subregion, None, sigma_level, loop_gain,
                          tolerance, accuracy, major_loop_miter, minor_loop_miter, conv_mode, enforce_positivity,
                          edge_suppression, edge_offset, flux_threshold, neg_comp, edge_excl, int_excl, precision)
        self._run(True, start_scale, stop_scale, prm)
        self.complete = False

    def restore(self):
        """Model convolved with the fitted clean beam plus the residual (pymoresane/main.py:601-616).

        The beam fit runs on the host (beam_fit); the convolution runs on the GPU:
        linear (2N frame) when the PSF is twice the image, circular otherwise.
        """
        from .beam_fit import beam_fit
        from .iuwt_convolution import fft_convolve, psf_spectrum
        clean_beam, beam_params = beam_fit(self.psf_data, self.cdelt1, self.cdelt2)
        n = self.dirty_data_shape[0]
        mode = "linear" if self.psf_data_shape[0] == 2 * n else "circular"
        conv = fft_convolve(self.model, psf_spectrum(clean_beam.astype(np.float32)), conv_mode=mode)
        self.restored = (conv.astype(np.float64) + self.residual).astype(np.float32)
        return beam_params

    def handle_input(self, input_hdr):
        """Index that keeps the RA/DEC axes of a FITS cube and takes plane 0 elsewhere (main.py:618-634)."""
        naxis = int(input_hdr["NAXIS"])
        idx = [0] * naxis
        for i in range(naxis):
            ctype = str(input_hdr.get("CTYPE%d" % (i + 1), ""))
            if ctype.startswith("RA"):
                idx[-1] = slice(None)
            if ctype.startswith("DEC"):
                idx[-2] = slice(None)
        return tuple(idx)

    def save_fits(self, data, name):
        """Write data as a (1, 1, N, N) float32 FITS with the dirty image's header (main.py:636-646)."""
        data = np.asarray(data)
        _fits.write(name, data.reshape(1, 1, data.shape[0], data.shape[1]), getattr(self, "img_hdr", None))

    def make_logger(self, level="INFO", log_file="PyMORESANE.log"):
        """This module's logger streaming to stderr and to log_file at `level` (main.py:648-676).

        Handlers added by an earlier call are replaced rather than stacked.
        """
        level = getattr(logging, level.upper())
        logger.setLevel(logging.DEBUG)
        for h in [h for h in logger.handlers if getattr(h, "_pms_made", False)]:
            logger.removeHandler(h)
            h.close()
        fmt = logging.Formatter("%(asctime)s [%(levelname)s]: %(message)s", datefmt="[%m/%d/%Y] [%I:%M:%S]")
        handlers = [logging.StreamHandler()]
        if log_file:
            handlers.append(logging.FileHandler(log_file, mode="w"))
        for h in handlers:
            h.setLevel(level)
            h.setFormatter(fmt)
            h._pms_made = True
            logger.addHandler(h)
        return logger


def _box5(a):
    """5 x 5 box sum with zero padding, same size (mask smoothing, main.py:711-715)."""
    p = np.pad(a, 2)
    out = np.zeros_like(a, dtype=np.float64)
    for dr in range(5):
        for dc in range(5):
            out += p[dr:dr + a.shape[0], dc:dc + a.shape[1]]
    return out


class FitsImage(DataImage):
    """DataImage read from FITS files (pymoresane/main.py:678-733)."""

    def __init__(self, image_name, psf_name, mask_name=None):
        self.image_name = image_name
        self.psf_name = psf_name
        self.img_hdr, img = _fits.read(image_name)
        self.psf_hdr, psf = _fits.read(psf_name)
        cdelt1 = self.psf_hdr.get("CDELT1")
        cdelt2 = self.psf_hdr.get("CDELT2", cdelt1)
        mask = None
        if mask_name is not None:
            _, m = _fits.read(mask_name)
            m = m.reshape(m.shape[-2], m.shape[-1]).astype(np.float64)
            m = m / np.max(m)
            m = _box5(m)
            mask = m / np.max(m)
        super(FitsImage, self).__init__(img[self.handle_input(self.img_hdr)], psf[self.handle_input(self.psf_hdr)],
                                        mask, cdelt1, cdelt2)
        self.mask_name = mask_name

    def restore(self):
        beam_params = super(FitsImage, self).restore()
        self.img_hdr.set("BMAJ", float(beam_params[0]))
        self.img_hdr.set("BMIN", float(beam_params[1]))
        self.img_hdr.set("BPA", float(beam_params[2]))
        return beam_params


def main(argv=None):
    """``runsane`` (pymoresane/main.py:736-789): deconvolve a dirty/PSF FITS pair on the GPU, by scale
    unless ``--singlerun``, and write the model, residual and restored images as FITS.

    Output names are ``OUTPUTNAME_model.fits``, ``_residual.fits`` and ``_restored.fits`` unless
    ``--modelname`` / ``--residualname`` / ``--restoredname`` give them; without OUTPUTNAME all three
    must be given (ValueError otherwise).  Returns the FitsImage.
    """
    import time
    from .parser import handle_parser
    args = handle_parser(argv)
    if args.outputname is None and (args.residualname is None or args.restoredname is None or
                                    args.modelname is None):
        raise ValueError("If outputname is unspecified, residualname, restoredname and modelname must be present.")
    data = FitsImage(args.dirty, args.psf, args.mask)
    log = data.make_logger(args.loglevel)
    log.info("Parameters:\n" + str(args)[10:-1])
    start = time.time()
    common = dict(subregion=args.subregion, sigma_level=args.sigmalevel, loop_gain=args.loopgain,
                  tolerance=args.tolerance, accuracy=args.accuracy, major_loop_miter=args.majorloopmiter,
                  minor_loop_miter=args.minorloopmiter, all_on_gpu=args.allongpu, decom_mode=args.decommode,
                  core_count=args.corecount, conv_device=args.convdevice, conv_mode=args.convmode,
                  extraction_mode=args.extractionmode, enforce_positivity=args.enforcepositivity,
                  edge_suppression=args.edgesuppression, edge_offset=args.edgeoffset,
                  flux_threshold=args.fluxthreshold, neg_comp=args.negcomp, edge_excl=args.edgeexcl,
                  int_excl=args.intexcl, precision=args.precision)
    if args.singlerun:
        data.moresane(scale_count=args.scalecount, **common)
    else:
        data.moresane_by_scale(start_scale=args.startscale, stop_scale=args.stopscale, **common)
    log.info("Elapsed time was %s." % time.strftime("%H:%M:%S", time.gmtime(time.time() - start)))
    data.save_fits(data.model, args.modelname or args.outputname + "_model.fits")
    data.save_fits(data.residual, args.residualname or args.outputname + "_residual.fits")
    data.restore()
    data.save_fits(data.restored, args.restoredname or args.outputname + "_restored.fits")
    return data


if __name__ == "__main__":
    main()
