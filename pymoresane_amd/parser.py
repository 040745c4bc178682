"""Command-line arguments of ``runsane`` (pymoresane/parser.py:4-125).

Same flags, meanings and defaults as the reference, except that the device modes
(``-dm``, ``-cd``, ``-em``) accept ``hip`` and default to it: this package runs
only the HIP path, and the reference's CPU choices (``ser``, ``mp``, ``cpu``)
are still parsed but then rejected by ``moresane`` with an error naming the HIP
path.  ``gpu`` is an alias of ``hip``.
"""

import argparse


def make_parser():
    p = argparse.ArgumentParser(
        description="Runs the MORESANE deconvolution on the GPU (HIP) with the specified arguments. In the event "
                    "that non-critical parameters are missing, the defaults will be used.",
        formatter_class=argparse.ArgumentDefaultsHelpFormatter)
    p.add_argument("dirty", help="File name and location of the input dirty map .fits file.")
    p.add_argument("psf", help="File name and location input psf .fits file.")
    p.add_argument("outputname", nargs="?", default=None,
                   help="File name and location of the output model and residual .fits files.")
    p.add_argument("-sr", "--singlerun", action="store_true",
                   help="Run in single-run mode instead of scale-by-scale mode.")
    p.add_argument("-sbr", "--subregion", default=None, type=int,
                   help="Pixel width of the central region of the dirty .fits which is to be deconvolved.")
    p.add_argument("-sc", "--scalecount", default=None, type=int,
                   help="Maximum number of wavelet scales (single-run mode only).")
    p.add_argument("-sts", "--startscale", default=1, type=int, help="First scale in the scale-by-scale case.")
    p.add_argument("-sps", "--stopscale", default=20, type=int, help="Last scale in the scale-by-scale case.")
    p.add_argument("-sl", "--sigmalevel", default=4, type=float,
                   help="Sigma level at which the wavelet coefficients are thresholded.")
    p.add_argument("-lg", "--loopgain", default=0.2, type=float, help="Loop gain for the deconvolution step.")
    p.add_argument("-tol", "--tolerance", default=0.75, type=float,
                   help="Fraction of the maximum wavelet coefficient used as the significance threshold.")
    p.add_argument("-ac", "--accuracy", default=1e-6, type=float,
                   help="Relative change in the STD of the residual that stops the major loop.")
    p.add_argument("-malm", "--majorloopmiter", default=100, type=int, help="Maximum major loop iterations.")
    p.add_argument("-milm", "--minorloopmiter", default=50, type=int, help="Maximum minor loop iterations.")
    p.add_argument("-aog", "--allongpu", action="store_true", help="Run everything on the GPU (always the case).")
    p.add_argument("-dm", "--decommode", default="hip", choices=["hip", "gpu", "ser", "mp"],
                   help="Device of the wavelet decompositions.")
    p.add_argument("-cc", "--corecount", default=1, type=int, help="Accepted for compatibility; unused.")
    p.add_argument("-cd", "--convdevice", default="hip", choices=["hip", "gpu", "cpu"],
                   help="Device of the convolutions.")
    p.add_argument("-cm", "--convmode", default="circular", choices=["circular", "linear"],
                   help="Circular or linear convolution.")
    p.add_argument("-em", "--extractionmode", default="hip", choices=["hip", "gpu", "cpu"],
                   help="Device of the source extraction.")
    p.add_argument("-ep", "--enforcepositivity", action="store_true",
                   help="Constrain the model to be strictly positive.")
    p.add_argument("-es", "--edgesuppression", action="store_true",
                   help="Suppress the edge-corrupted wavelet coefficients.")
    p.add_argument("-eo", "--edgeoffset", default=0, type=int, help="Additional offset along the edges.")
    p.add_argument("-ll", "--loglevel", default="INFO", choices=["DEBUG", "INFO", "WARNING", "ERROR", "CRITICAL"],
                   help="Logging level.")
    p.add_argument("-m", "--mask", default=None, help="File name and location of the input .fits mask.")
    p.add_argument("-ft", "--fluxthreshold", default=0, type=float,
                   help="Flux threshold level for shallow deconvolution.")
    p.add_argument("-rn", "--residualname", default=None, help="Specific residual image name.")
    p.add_argument("-mn", "--modelname", default=None, help="Specific model image name.")
    p.add_argument("-rsn", "--restoredname", default=None, help="Specific restored image name.")
    p.add_argument("-nc", "--negcomp", action="store_true",
                   help="Use anticorrelations in the source extraction step.")
    p.add_argument("-ee", "--edgeexcl", type=int, default=0,
                   help="Pixels to exclude from the edges when estimating the noise.")
    p.add_argument("-ie", "--intexcl", type=int, default=0,
                   help="Pixels to exclude from the centre when estimating the noise.")
    # not in the reference: the arithmetic of the device path
    p.add_argument("-pr", "--precision", default="float32", choices=["float32", "float64"],
                   help="Bulk arithmetic of the HIP path: float32 (float64 inner products and step sizes) or "
                        "float64 throughout.")
    return p


def handle_parser(argv=None):
    """Parsed arguments (pymoresane/parser.py:4); ``argv`` defaults to the command line."""
    return make_parser().parse_args(argv)
