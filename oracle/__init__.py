"""CPU float64 oracle for the MORESANE hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in plain numpy/scipy at float64, what the reference
PyMORESANE 0.3.6 computes on the path that the HIP library accelerates:

* ``iuwt64``        -- iuwt.py:66-191 (serial a trous IUWT)
* ``mp_iuwt``       -- iuwt.py:193-382 (fork-per-pass strip IUWT, CPU timing only)
* ``convolution64`` -- iuwt_convolution.py:18-73, 142-158 (FFT convolution)
* ``toolbox64``     -- iuwt_toolbox.py:17-154, 292-306 (MAD threshold, extraction, SNR)
* ``moresane64``    -- main.py:53-599 (major loop, CG minor loop, by-scale loop)

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker (or the timed
CPU baseline).  The product package ``pymoresane_amd`` never imports it and
has no CPU fallback.

How each part is pinned (the reference cannot be imported: it is Python 2 and
needs astropy/pycuda, and it ships no golden vectors):

* a_trous / IUWT -- equal to ``scipy.ndimage.correlate1d`` with a dilated B3
  kernel and ``mode='reflect'`` on both axes; perfect reconstruction; constant
  image gives zero details; brute-force 2-D direct sum on tiny images.
* fft_convolve   -- equal to ``scipy.signal.fftconvolve(x, psf)[N:2N, N:2N]``
  (linear) and to a brute-force circular sum (circular).
* estimate_threshold / apply_threshold / snr_ratio -- hand-built inputs with
  known medians, values exactly on the threshold, closed-form norms.
* source_extraction -- ``scipy.ndimage.label`` checked against a brute-force
  BFS labeller; hand-built nested multi-scale cases for the keep rule.
* moresane / moresane_by_scale -- every sub-operation is pinned as above;
  the end-to-end trajectory is pinned only by invariants (residual identity,
  positivity, trace consistency).  The reference's own mixed-precision
  trajectory (float32 model, Python 2 casting) is parity unpinned.
"""
