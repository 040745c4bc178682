"""Argument checks of the reference-shaped API that run before any device work (CPU).

The reference raises ValueError for an image of odd side before it does anything
(pymoresane/main.py:101-103); this package does the same, rejects the shapes its device driver does not
take (non-square images, a PSF that is neither the image size nor twice it), and rejects the
reference's CPU modes and unknown modes instead of returning None (pymoresane/iuwt.py:36-41,
iuwt_toolbox.py:86-94).  None of these calls touches the GPU, so they run here.
"""

import numpy as np
import pytest

from pymoresane_amd import iuwt, iuwt_convolution as conv, iuwt_toolbox as tools
from pymoresane_amd.main import DataImage


def image(n, psf_n=None):
    return DataImage(np.zeros((n, n), np.float32), np.zeros((psf_n or 2 * n,) * 2, np.float32), None, 1.0, 1.0)


@pytest.mark.parametrize("by_scale", [False, True])
def test_odd_image_side_is_rejected_as_the_reference_does(by_scale):
    img = image(63, 126)
    run = img.moresane_by_scale if by_scale else img.moresane
    with pytest.raises(ValueError, match="Image size is uneven"):
        run()


def test_shapes_the_driver_does_not_take_are_rejected():
    with pytest.raises(ValueError, match="square"):
        DataImage(np.zeros((64, 32), np.float32), np.zeros((128, 128), np.float32), None, 1.0, 1.0).moresane()
    with pytest.raises(ValueError, match="PSF"):
        image(64, 96).moresane()


def test_invalid_options_are_rejected():
    with pytest.raises(ValueError, match="conv_mode"):
        image(64).moresane(conv_mode="full")
    with pytest.raises(ValueError, match="precision"):
        image(64).moresane(precision="float16")
    with pytest.raises(ValueError, match="CPU path"):
        image(64).moresane(decom_mode="ser")
    with pytest.raises(ValueError, match="unknown"):
        image(64).moresane(extraction_mode="opencl")


@pytest.mark.parametrize("mode", ["ser", "mp", "cpu", "cuda"])
def test_operations_reject_modes_they_do_not_run(mode):
    x = np.zeros((32, 32), np.float32)
    with pytest.raises(ValueError):
        iuwt.iuwt_decomposition(x, 3, mode=mode)
    with pytest.raises(ValueError):
        iuwt.iuwt_recomposition(np.zeros((3, 32, 32), np.float32), mode=mode)
    with pytest.raises(ValueError):
        tools.source_extraction(np.zeros((2, 32, 32), np.float32), 0.75, mode=mode)
    with pytest.raises(ValueError):
        conv.fft_convolve(x, np.zeros((64, 33), np.complex64), conv_device=mode)
