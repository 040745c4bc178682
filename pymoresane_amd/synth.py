"""Seeded synthetic dirty/PSF generators for the MORESANE workloads.

These build inputs only; they are not part of the deconvolution path.  Every
choice is fixed here so that a seed reproduces the same float32 arrays:

* PSF: an elliptical Gaussian core (major/minor FWHM, position angle measured
  from the +row axis) centred at ``P/2``, plus a ring ``amp * exp(-(r-r0)^2 /
  (2 * 1.5^2))`` with ``r0 = 12`` px standing in for an Airy-like sidelobe,
  normalised to a peak of 1.
* Sky: single-pixel point sources at uniform integer positions with
  log-uniform flux, plus circular Gaussian blobs with uniform sigma and
  amplitude.
* Dirty image: the linear convolution of the sky with the 2N PSF (the
  algorithm's own forward operator, ``out[i] = sum_u sky[u] psf[i-u+N]``),
  computed in float64, plus white Gaussian noise, cast to float32.
"""

import numpy as np


def fwhm_to_sigma(fwhm):
    return fwhm / (2.0 * np.sqrt(2.0 * np.log(2.0)))


def make_psf(p, fwhm_major=4.0, fwhm_minor=3.0, pa_deg=30.0, ring_amp=0.05, ring_r0=12.0, ring_w=1.5):
    """P x P float32 PSF with its centre at (P/2, P/2), peak 1."""
    c = p // 2
    y = (np.arange(p) - c).astype(np.float64)
    rr, cc = np.meshgrid(y, y, indexing="ij")
    th = np.deg2rad(pa_deg)
    a = rr * np.cos(th) + cc * np.sin(th)
    b = -rr * np.sin(th) + cc * np.cos(th)
    sa, sb = fwhm_to_sigma(fwhm_major), fwhm_to_sigma(fwhm_minor)
    psf = np.exp(-0.5 * ((a / sa) ** 2 + (b / sb) ** 2))
    if ring_amp:
        r = np.sqrt(rr ** 2 + cc ** 2)
        psf += ring_amp * np.exp(-((r - ring_r0) ** 2) / (2.0 * ring_w ** 2))
    psf /= psf.max()
    return psf.astype(np.float32)


def make_gaussian_psf(p, fwhm):
    """Separable circular Gaussian P x P PSF centred at P/2 (peak 1), built without a P^2 float64 temp."""
    c = p // 2
    g = np.exp(-0.5 * ((np.arange(p) - c) / fwhm_to_sigma(fwhm)) ** 2).astype(np.float32)
    return np.outer(g, g)


def make_sky(n, rng, n_points=40, flux_lo=-3.0, flux_hi=0.0, n_blobs=4, sig_lo=2.0, sig_hi=8.0,
             amp_lo=0.05, amp_hi=0.3):
    sky = np.zeros((n, n))
    rows = rng.integers(0, n, n_points)
    cols = rng.integers(0, n, n_points)
    flux = 10.0 ** rng.uniform(flux_lo, flux_hi, n_points)
    np.add.at(sky, (rows, cols), flux)
    y = np.arange(n, dtype=np.float64)
    for _ in range(n_blobs):
        cy, cx = rng.uniform(0, n, 2)
        sig = rng.uniform(sig_lo, sig_hi)
        amp = rng.uniform(amp_lo, amp_hi)
        gy = np.exp(-0.5 * ((y - cy) / sig) ** 2)
        gx = np.exp(-0.5 * ((y - cx) / sig) ** 2)
        sky += amp * np.outer(gy, gx)
    return sky


def linear_convolve(sky, psf):
    """``out[i, j] = sum_uv sky[u, v] psf[i-u+N, j-v+N]`` for an N x N sky (float64).

    A circular convolution of size max(P, 2N) has no wrap-around on output
    rows/columns [N, 2N), so that window is the linear result.
    """
    from scipy import fft as sfft
    n = sky.shape[0]
    p = psf.shape[0]
    size = max(p, 2 * n)
    a = np.zeros((size, size))
    a[:n, :n] = sky
    b = np.zeros((size, size))
    b[:p, :p] = psf
    spec = sfft.rfft2(a, workers=-1) * sfft.rfft2(b, workers=-1)
    full = sfft.irfft2(spec, s=(size, size), workers=-1)
    return full[n:2 * n, n:2 * n]


def make_pair(n, seed, psf_factor=2, fwhm_major=4.0, fwhm_minor=3.0, pa_deg=30.0, n_points=40,
              n_blobs=4, noise=1e-3, ring_amp=0.05):
    """(dirty float32 N x N, psf float32 2N x 2N) drawn by the seeded generator."""
    rng = np.random.default_rng(seed)
    psf = make_psf(psf_factor * n, fwhm_major, fwhm_minor, pa_deg, ring_amp=ring_amp)
    sky = make_sky(n, rng, n_points=n_points, n_blobs=n_blobs)
    dirty = linear_convolve(sky, psf) + rng.normal(0.0, noise, (n, n))
    return dirty.astype(np.float32), psf


def config_c0(seed=0):
    """256^2 dirty, 512^2 PSF, 40 points + 4 blobs, noise 1e-3."""
    return make_pair(256, seed)


def config_c2(seed=2, n=2048):
    """2048^2 dirty, 4096^2 PSF, 2000 points + 50 blobs, noise 1e-4."""
    return make_pair(n, seed, n_points=2000, n_blobs=50, noise=1e-4)


def config_c1(seed=1, n=4096):
    """4096^2 image N(0,1) + 200 points (flux 10^U(-2,1)) and a 2N Gaussian PSF of FWHM 6 px."""
    rng = np.random.default_rng(seed)
    img = rng.standard_normal((n, n), dtype=np.float32)
    rows = rng.integers(0, n, 200)
    cols = rng.integers(0, n, 200)
    flux = (10.0 ** rng.uniform(-2.0, 1.0, 200)).astype(np.float32)
    np.add.at(img, (rows, cols), flux)
    return img, make_gaussian_psf(2 * n, 6.0)


BATCH_SIZES = ((256, 32), (512, 32), (1024, 24), (2048, 8))


def config_c3_pair(i, side):
    """Pair i of the ragged batch: seed 3000+i, PSF FWHM ~ U(3, 8) px."""
    rng = np.random.default_rng(3000 + i)
    fwhm = rng.uniform(3.0, 8.0)
    return make_pair(side, 3000 + i, fwhm_major=fwhm, fwhm_minor=0.75 * fwhm)


def config_c3_sides():
    out = []
    for side, count in BATCH_SIZES:
        out.extend([side] * count)
    return out


SERVE_SIDES = (512, 1024, 2048, 4096)
SERVE_PROBS = (0.5, 0.3, 0.15, 0.05)


def config_c4_requests(count, rate=20.0, seed=4):
    """(arrival_time_s, side, seed) for ``count`` Poisson arrivals at ``rate`` per second."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, count)
    sides = rng.choice(SERVE_SIDES, size=count, p=SERVE_PROBS)
    t = np.cumsum(gaps)
    return [(float(t[k]), int(sides[k]), 4000 + k) for k in range(count)]
