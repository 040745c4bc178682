"""runsane command line (pymoresane/parser.py, pymoresane/main.py:736-789).

CPU: the parser keeps the reference's flags and defaults, and main() hands them to
FitsImage.moresane / moresane_by_scale and writes the three FITS outputs (with a
stand-in FitsImage, no GPU).  GPU: the reference's own end-to-end case
(tests/Makefile:8, ``runsane -ep dirty beam test_output``) through the command
line, against the oracle's fixture.
"""

import os
import sys

import numpy as np
import pytest

from pymoresane_amd import fits, main as pmain
from pymoresane_amd.parser import handle_parser

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_parser_defaults_follow_the_reference():
    a = handle_parser(["d.fits", "p.fits"])
    # pymoresane/parser.py:14-123 defaults; device modes default to the HIP path
    assert a.outputname is None and not a.singlerun and a.subregion is None and a.scalecount is None
    assert (a.startscale, a.stopscale, a.sigmalevel, a.loopgain, a.tolerance, a.accuracy) == (1, 20, 4, 0.2, 0.75,
                                                                                              1e-6)
    assert (a.majorloopmiter, a.minorloopmiter, a.corecount, a.edgeoffset, a.fluxthreshold) == (100, 50, 1, 0, 0)
    assert (a.convmode, a.loglevel, a.edgeexcl, a.intexcl) == ("circular", "INFO", 0, 0)
    assert (a.decommode, a.convdevice, a.extractionmode) == ("hip", "hip", "hip")
    assert not (a.allongpu or a.enforcepositivity or a.edgesuppression or a.negcomp)
    b = handle_parser(["d.fits", "p.fits", "out", "-sr", "-sc", "5", "-ep", "-nc", "-cm", "linear", "-dm", "gpu",
                       "-ee", "3", "-ie", "4", "-sbr", "128", "-ft", "0.5", "-milm", "30"])
    assert b.singlerun and b.scalecount == 5 and b.enforcepositivity and b.negcomp and b.convmode == "linear"
    assert b.decommode == "gpu" and (b.edgeexcl, b.intexcl, b.subregion, b.fluxthreshold) == (3, 4, 128, 0.5)
    assert b.minorloopmiter == 30
    assert a.precision == "float32" and handle_parser(["d.fits", "p.fits", "-pr", "float64"]).precision == "float64"
    with pytest.raises(SystemExit):
        handle_parser(["d.fits", "p.fits", "-cm", "spherical"])
    with pytest.raises(SystemExit):
        handle_parser(["d.fits", "p.fits", "-pr", "float16"])


def test_main_requires_output_names():
    with pytest.raises(ValueError, match="residualname, restoredname and modelname"):
        pmain.main(["d.fits", "p.fits", "-mn", "m.fits"])


class _FakeImage(object):
    """Stands in for FitsImage: records the call, returns fixed planes, writes FITS for real."""

    calls = []

    def __init__(self, dirty, psf, mask=None):
        self.names = (dirty, psf, mask)
        self.img_hdr = fits.Header([("CTYPE1", "RA---TAN", ""), ("CTYPE2", "DEC--TAN", "")])

    def make_logger(self, level="INFO"):
        return pmain.DataImage.make_logger(self, level, log_file=None)

    def _done(self, kind, kw):
        _FakeImage.calls.append((kind, self.names, kw))
        self.model = np.full((8, 8), 1.0, np.float32)
        self.residual = np.full((8, 8), 2.0, np.float32)

    def moresane(self, **kw):
        self._done("moresane", kw)

    def moresane_by_scale(self, **kw):
        self._done("by_scale", kw)

    def restore(self):
        self.restored = self.model + self.residual

    save_fits = pmain.DataImage.save_fits


def test_main_forwards_arguments_and_writes_outputs(tmp_path, monkeypatch):
    monkeypatch.setattr(pmain, "FitsImage", _FakeImage)
    _FakeImage.calls = []
    out = str(tmp_path / "run")
    pmain.main(["d.fits", "p.fits", out, "-ep", "-lg", "0.3", "-sps", "6", "-m", "mask.fits", "-ll", "WARNING"])
    kind, names, kw = _FakeImage.calls[-1]
    assert kind == "by_scale" and names == ("d.fits", "p.fits", "mask.fits")
    assert kw["start_scale"] == 1 and kw["stop_scale"] == 6 and kw["loop_gain"] == 0.3
    assert kw["enforce_positivity"] and kw["conv_mode"] == "circular" and kw["minor_loop_miter"] == 50
    assert kw["decom_mode"] == kw["conv_device"] == kw["extraction_mode"] == "hip"
    assert kw["precision"] == "float32"
    for suffix, val in (("_model.fits", 1.0), ("_residual.fits", 2.0), ("_restored.fits", 3.0)):
        _, a = fits.read(out + suffix)
        assert a.shape == (1, 1, 8, 8) and np.all(a == val)
    # single run with explicit names
    names = [str(tmp_path / n) for n in ("m.fits", "r.fits", "s.fits")]
    pmain.main(["d.fits", "p.fits", "-sr", "-sc", "4", "-mn", names[0], "-rn", names[1], "-rsn", names[2],
                "-ll", "WARNING"])
    kind, _, kw = _FakeImage.calls[-1]
    assert kind == "moresane" and kw["scale_count"] == 4 and not kw["enforce_positivity"]
    assert all(os.path.exists(n) for n in names)


@pytest.mark.gpu
def test_runsane_tart_end_to_end(hip, tmp_path, monkeypatch):
    """runsane -ep on the TART pair (tests/Makefile:8): model and residual within 1e-4 of the oracle's
    by-scale run, the same decision trace, and the three FITS files written."""
    print("runsane -ep TART"); sys.stdout.flush()
    g = np.load(os.path.join(GOLD, "tart_inputs.npz"), allow_pickle=False)
    z = np.load(os.path.join(GOLD, "tart_moresane.npz"), allow_pickle=False)
    h = fits.Header([("CTYPE1", "RA---TAN", ""), ("CTYPE2", "DEC--TAN", ""), ("CDELT1", 0.703125, ""),
                     ("CDELT2", 0.703125, "")])
    dp, bp = str(tmp_path / "dirty.fits"), str(tmp_path / "beam.fits")
    fits.write(dp, g["dirty"], h)
    fits.write(bp, g["beam"], h)
    monkeypatch.chdir(tmp_path)  # PyMORESANE.log lands here
    out = str(tmp_path / "test_output")
    img = pmain.main([dp, bp, out, "-ep"])

    def rel(a, b):
        a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    from test_golden import discrete, trace_of
    assert discrete(img.trace) == discrete(trace_of(z))
    _, model = fits.read(out + "_model.fits")
    _, resid = fits.read(out + "_residual.fits")
    _, restored = fits.read(out + "_restored.fits")
    assert rel(model.reshape(256, 256), z["model"]) < 1e-4
    assert rel(resid.reshape(256, 256), z["residual"]) < 1e-4
    np.testing.assert_array_equal(restored.reshape(256, 256), img.restored)
    assert os.path.exists(str(tmp_path / "PyMORESANE.log"))
    hdr, _ = fits.read(out + "_restored.fits")
    assert "BMAJ" in hdr


def test_fitsimage_reads_cubes_and_smooths_the_mask(tmp_path):
    """FitsImage (main.py:678-733) keeps the RA/DEC plane of (1, 1, N, N) cubes, takes CDELT1 from the PSF
    header, and turns a mask FITS into mask/max, a 5x5 box filter ('same'), then /max again
    (main.py:711-715, with its use-before-assignment fixed) -- here against scipy's fftconvolve."""
    from scipy.signal import fftconvolve
    rng = np.random.default_rng(8)
    n = 32
    h = fits.Header([("CTYPE1", "RA---SIN", ""), ("CTYPE2", "DEC--SIN", ""), ("CTYPE3", "FREQ", ""),
                     ("CTYPE4", "STOKES", ""), ("CDELT1", 0.25, ""), ("CDELT2", 0.25, "")])
    dirty = rng.standard_normal((1, 1, n, n)).astype(np.float32)
    psf = rng.standard_normal((1, 1, 2 * n, 2 * n)).astype(np.float32)
    mask = (rng.random((1, 1, n, n)) > 0.7).astype(np.float32) * 3.0
    paths = [str(tmp_path / f) for f in ("d.fits", "p.fits", "m.fits")]
    for p, a in zip(paths, (dirty, psf, mask)):
        fits.write(p, a, h)
    img = pmain.FitsImage(*paths)
    np.testing.assert_array_equal(img.dirty_data, dirty[0, 0])
    np.testing.assert_array_equal(img.psf_data, psf[0, 0])
    assert img.cdelt1 == 0.25 and img.mask_name == paths[2]
    m = mask[0, 0].astype(np.float64) / mask.max()
    m = fftconvolve(m, np.ones((5, 5)), mode="same")
    np.testing.assert_allclose(img.mask, (m / m.max()).astype(np.float32), rtol=0, atol=1e-6)
    assert pmain.FitsImage(*paths[:2]).mask is None
