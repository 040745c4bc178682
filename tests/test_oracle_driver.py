"""Pins of the oracle's MORESANE driver (oracle/moresane64.py) that do not reuse its own formulas.

* Control flow of the minor loop (main.py:427-475): the SNR is scripted (the oracle's
  ``snr_ratio`` is replaced by a fixed sequence), so each exit rule of the reference is reached on
  purpose and the expected outcome -- exit code, iteration count, min_scale change, and which CG
  iterate becomes the model (main.py:441, 456-457) -- is written down by hand from the reference's
  rules, not computed by the oracle.
* Invariants of a real run: the residual equals the dirty image minus the linear convolution of
  the model computed by ``scipy.signal.fftconvolve`` (main.py:485, closed form of the linear
  mode); the model is non-negative under positivity; the first scale choice equals the argmax of
  normalised maxima computed here with scipy's ``correlate1d`` a trous, ``np.median`` and the
  strict threshold (main.py:278-306); ``scale_adjust`` is the scale above the highest empty
  scale below the maximum (main.py:321-335).
"""

import numpy as np
import pytest
from scipy import ndimage, signal

from oracle import moresane64 as om
from pymoresane_amd import synth

B3 = np.array([1.0, 4.0, 6.0, 4.0, 1.0]) / 16.0


@pytest.fixture(scope="module")
def pair():
    # seed 7: the maximum normalised coefficient sits at the coarsest of three scales
    return synth.make_pair(32, 7, n_points=8, n_blobs=1, noise=1e-3)


def scripted(monkeypatch, values):
    """Replace the SNR with ``values`` in order, then 45 dB (a first-iteration false detection)."""
    seq = list(values)
    seen = []

    def fake_snr(a, b):
        seen.append(np.array(b))
        return float(seq.pop(0)) if seq else 45.0

    monkeypatch.setattr(om.tools, "snr_ratio", fake_snr)
    return seen


def run(pair, **kw):
    dirty, psf = pair
    img = om.OracleImage(dirty, psf)
    args = dict(scale_count=3, loop_gain=1.0, major_loop_miter=1, minor_loop_miter=30, enforce_positivity=False)
    args.update(kw)
    img.moresane(**args)
    return img


def minors(trace):
    return [(t[2], t[3]) for t in trace if t[0] == "minor"]


def test_fell_ok_keeps_the_previous_iterate(monkeypatch, pair):
    # iteration 3 lowers the SNR (24 <= 25) above 10.5: exit 3, min_scale 0, model from x_2
    scripted(monkeypatch, [20, 25, 24])
    fell = run(pair)
    assert minors(fell.trace) == [(3, om.EXIT_FELL_OK)]
    assert [t[0] for t in fell.trace] == ["scales", "minor", "update"]
    # accept at iteration 2 (SNR > 40): the model is x_2 itself
    scripted(monkeypatch, [20, 45])
    acc2 = run(pair)
    assert minors(acc2.trace) == [(2, om.EXIT_ACCEPT)]
    # accept at iteration 3: the model is x_3, a different iterate
    scripted(monkeypatch, [20, 25, 45])
    acc3 = run(pair)
    assert minors(acc3.trace) == [(3, om.EXIT_ACCEPT)]
    np.testing.assert_array_equal(fell.model, acc2.model)
    assert np.abs(fell.model - acc3.model).max() > 1e-6 * np.abs(acc3.model).max()


def test_snr_falling_at_iteration_two_does_not_exit(monkeypatch, pair):
    # the fall rule needs more than two iterations (main.py:446): 19 <= 20 at iteration 2 goes on
    scripted(monkeypatch, [20, 19, 18])
    img = run(pair)
    assert minors(img.trace) == [(3, om.EXIT_FELL_OK)]


def test_fell_low_raises_min_scale_until_all_scales_fail(monkeypatch, pair):
    # exit 4 (SNR fell at or below 10.5) moves min_scale to 1; the later loops see 45 dB at their
    # first iteration (false detection, exit 1) and move it on to scale_count: "poor", no update
    scripted(monkeypatch, [5, 6, 6])
    img = run(pair)
    kinds = [t[0] for t in img.trace]
    assert kinds == ["scales", "minor", "scales", "minor", "scales", "minor", "poor"]
    assert minors(img.trace) == [(3, om.EXIT_FELL_LOW), (1, om.EXIT_FIRST_HIGH), (1, om.EXIT_FIRST_HIGH)]
    assert [t[2] for t in img.trace if t[0] == "scales"] == [0, 1, 2]
    assert not img.model.any()
    # the decomposition is not redone while min_scale > 0 (main.py:280): each later choice is the
    # argmax over the same maxima restricted to scales >= min_scale, so max_scale never drops
    # below min_scale + 1
    for t in img.trace:
        if t[0] == "scales":
            assert t[3] >= t[2] + 1


def test_iteration_cap_keeps_the_last_iterate(monkeypatch, pair):
    # cap (exit 0) with a final SNR above 10.5: min_scale 0 and the model is x_4 (main.py:456-457
    # assign x = xn at the end of every iteration that did not exit)
    scripted(monkeypatch, [11, 12, 13, 14])
    cap = run(pair, minor_loop_miter=4)
    assert minors(cap.trace) == [(4, om.EXIT_CAP)]
    scripted(monkeypatch, [11, 12, 13, 45])
    acc = run(pair, minor_loop_miter=4)
    assert minors(acc.trace) == [(4, om.EXIT_ACCEPT)]
    np.testing.assert_array_equal(cap.model, acc.model)
    # cap with a final SNR at or below 10.5 and min_scale 0: the update still happens (main.py:466)
    scripted(monkeypatch, [5, 6, 7, 8])
    low = run(pair, minor_loop_miter=4)
    assert minors(low.trace) == [(4, om.EXIT_CAP)]
    np.testing.assert_array_equal(low.model, acc.model)


def test_cap_with_low_snr_above_scale_zero_raises(monkeypatch, pair):
    # first loop: false detection (min_scale 1); second loop: cap at SNR 8 <= 10.5 with min_scale 1.
    # The reference would repeat that loop forever (nothing changes); the oracle raises instead.
    scripted(monkeypatch, [45, 5, 6, 7, 8])
    with pytest.raises(RuntimeError, match="loops forever"):
        run(pair, minor_loop_miter=4)


def test_first_iteration_snr_above_40_is_a_false_detection(monkeypatch, pair):
    scripted(monkeypatch, [45, 20, 45])
    img = run(pair)
    assert minors(img.trace)[:2] == [(1, om.EXIT_FIRST_HIGH), (2, om.EXIT_ACCEPT)]
    assert [t[2] for t in img.trace if t[0] == "scales"] == [0, 1]


def decomposition(x, scales):
    """Detail planes by scipy correlate1d (reflect) -- independent of oracle/iuwt64.py."""
    c0 = np.asarray(x, dtype=np.float64)
    out = []
    for s in range(scales):
        d = 1 << s
        k = np.zeros(4 * d + 1)
        k[::d] = B3
        smooth = lambda a: ndimage.correlate1d(ndimage.correlate1d(a, k, axis=0, mode="reflect"), k, axis=1,
                                               mode="reflect")
        c = smooth(c0)
        out.append(c0 - smooth(c))
        c0 = c
    return np.array(out)


def test_first_scale_choice_is_the_argmax_of_normalised_maxima(pair):
    dirty, psf = (np.asarray(a, dtype=np.float64) for a in pair)
    n, S = dirty.shape[0], 3
    img = run(pair, scale_count=S)
    first = next(t for t in img.trace if t[0] == "scales")
    dec = decomposition(dirty, S)
    thr = np.array([np.median(np.abs(p)) / 0.6745 for p in dec])
    kept = np.where(np.abs(dec) > 4 * thr[:, None, None], dec, 0.0)
    lo = psf.shape[0] // 2 - n // 2
    energy = np.sqrt((decomposition(psf[lo:lo + n, lo:lo + n], S) ** 2).sum(axis=(1, 2)))
    nsm = kept.max(axis=(1, 2)) / energy
    assert first[2] == 0
    assert first[3] == int(np.argmax(nsm)) + 1
    assert first[5] == pytest.approx(nsm.max(), rel=1e-12)


def test_scale_adjust_skips_below_the_highest_empty_scale(monkeypatch, pair):
    # empty thresholded scale 1 and a damped scale 0: the maximum sits at scale >= 2, so the
    # scales at and below the empty one are ignored (scale_adjust = 2)
    real = om.tools.apply_threshold

    def emptied(dec, thr, sigma_level=4):
        out = real(dec, thr, sigma_level)
        out[1] = 0.0
        out[0] *= 1e-6
        return out

    monkeypatch.setattr(om.tools, "apply_threshold", emptied)
    img = run(pair, scale_count=4, major_loop_miter=1)
    first = next(t for t in img.trace if t[0] == "scales")
    assert first[3] >= 3
    assert first[4] == 2


@pytest.mark.parametrize("by_scale", [False, True])
def test_residual_identity_and_positivity(by_scale):
    dirty, psf = synth.make_pair(64, 9, n_points=12, n_blobs=2, noise=1e-3)
    img = om.OracleImage(dirty, psf)
    kw = dict(loop_gain=0.2, major_loop_miter=6, enforce_positivity=True)
    if by_scale:
        img.moresane_by_scale(start_scale=1, stop_scale=3, **kw)
        calls = [t[1] for t in img.trace if t[0] == "byscale"]
        assert calls == list(range(1, 1 + len(calls))) and calls[-1] <= 3
    else:
        img.moresane(scale_count=3, **kw)
    assert any(t[0] == "update" for t in img.trace)
    n = dirty.shape[0]
    expect = dirty.astype(np.float64) - signal.fftconvolve(img.model, psf.astype(np.float64))[n:2 * n, n:2 * n]
    np.testing.assert_allclose(img.residual, expect, rtol=0, atol=1e-12 * np.abs(dirty).max())
    assert img.model.min() >= 0.0
    assert img.model.max() > 0.0
    last = [t for t in img.trace if t[0] == "update"][-1]
    if not by_scale:
        assert last[2] == pytest.approx(float(np.std(img.residual)), rel=1e-12)
