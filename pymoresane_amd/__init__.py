"""MI355X-native (gfx950 HIP) MORESANE deconvolution hot path.

Modules mirror the reference package (ratt-ru/PyMORESANE 0.3.6):
``iuwt``, ``iuwt_convolution``, ``iuwt_toolbox`` and ``main`` (DataImage /
FitsImage), with ``'hip'`` as the device mode.  All compute runs in
libpmship.so (C ABI in include/pymoresane_hip.h); there is no CPU fallback.
"""

__version__ = "0.1.0"
