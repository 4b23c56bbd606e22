"""Pins of the oracle's whole-image LSQR loop (fit_image_lsqr) on the
single-galaxy scene: a 64x64 image, ExpGalaxy flux 1000, r_e 3 px, 31x31
two-Gaussian pixelized PSF.

- the no-modelMask extent rule: halfsize 22 (the PSF radius hypot(15.5,
  15.5) = 21.9 dominates 4 r_e = 12), pixels floor(x - 22) .. ceil(x + 23)
  inclusive, clipped to the image, FFT size 64;
- with position parameters in pixels (cd^-1 = I) and noiseless data the loop
  walks back to the truth, the exact minimum (chi^2 = 0): a wrong sign or a
  swapped row of the position derivatives, a wrong shape step or a broken
  line search stops it elsewhere;
- with the reference's NullWCS (cd^-1 = 3600 I, hazard H1) the column-scaled
  position steps are 3600 times too small, so the position stays at its
  start while flux and shape move to their conditional optimum;
- a step gaining at most dchisq ends the loop (lsqr_optimizer.py:317-326).
"""
import math

import numpy as np

import oracle as O
from tractor_amd import workloads as W


def _scene():
    sc = W.config1_scene()
    return sc, np.ones((64, 64), np.float32)


def test_extent_rule_and_fft_size():
    sc, _ = _scene()
    t = sc['truth']
    x0, y0, p = O.image_galaxy_unit_patch(1, t, sc['psf'], (64, 64))
    assert O.galaxy_patch_halfsize(4 * 3., t[0], t[1], (64, 64), math.hypot(15.5, 15.5)) == 22
    assert (x0, y0) == (int(math.floor(t[0] - 22)), int(math.floor(t[1] - 22)))
    # inclusive range (get_overlapping_region, patch.pyx:20-66)
    assert p.shape == (int(math.ceil(t[1] + 23)) - y0 + 1, int(math.ceil(t[0] + 23)) - x0 + 1)
    assert O.fft_size(22) == 64
    # unit flux, nearly all of it inside the 47 x 47 patch
    assert 0.99 < p.sum() < 1.0 + 1e-6


def test_noiseless_fit_returns_to_truth_with_pixel_positions():
    sc, inv = _scene()
    t = sc['truth']
    data = O.image_model(1, t, sc['psf'], (64, 64))
    p, step, chi2, trace = O.fit_image_lsqr(1, sc['start'], sc['psf'], data, inv, np.eye(2),
                                            dchisq=1e-6, steps=50)
    assert chi2 < 1e-3, (chi2, trace)
    np.testing.assert_allclose(p[:2], t[:2], atol=1e-3)
    np.testing.assert_allclose(p[2], t[2], rtol=1e-4)
    np.testing.assert_allclose(p[3:], t[3:], atol=1e-3)
    # every accepted step improved, the last gained at most dchisq
    assert all(d > 0 for d, a in trace[:-1]) and trace[-1][0] <= 1e-6
    assert step == len(trace) - 1


def test_nullwcs_position_steps_are_3600x_too_small():
    """With noise: pixel positions converge in a handful of steps; with
    NullWCS's cd^-1 = 3600 I each step moves the position 3600 times less
    (about 4e-4 px at alpha = 2), every step still gains more than dchisq,
    and the loop runs to its 50-step limit with the position near its start
    and chi^2 well above the pixel-position fit's."""
    sc, inv = _scene()
    data = O.image_model(1, sc['truth'], sc['psf'], (64, 64)) + sc['noise']
    pp, sp, chip, trp = O.fit_image_lsqr(1, sc['start'], sc['psf'], data, inv, np.eye(2),
                                         dchisq=1e-3)
    assert sp < 10 and trp[-1][0] <= 1e-3
    np.testing.assert_allclose(pp[:2], sc['truth'][:2], atol=0.2)
    p, step, chi2, trace = O.fit_image_lsqr(1, sc['start'], sc['psf'], data, inv,
                                            np.eye(2) * 3600., dchisq=1e-3)
    assert step == 49 and all(d > 1e-3 for d, a in trace)
    moved = np.abs(p[:2] - sc['start'][:2])
    off = np.abs(sc['start'][:2] - sc['truth'][:2])
    assert np.all(moved < 0.1 * off), (moved, off)
    # about 1e-3 px or less per step at the largest alpha, 2
    assert np.all(moved / len(trace) < 1e-3)
    assert max(a for d, a in trace) == 2.
    assert chi2 > chip + 100.
