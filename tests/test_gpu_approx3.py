"""Ring-growth MoG evaluation on the device (trh_render_mog_approx3_batch)
against the oracle's restatement of c_gauss_2d_approx3 (tractor/approx3.c):
the set of evaluated pixels and the bbox must match exactly, values and
derivatives to FP64 rounding (1e-12 relative)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module', autouse=True)
def _need_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')


def _mix(K, seed):
    rng = np.random.RandomState(seed)
    amp = rng.uniform(0.2, 2., K)
    mean = rng.uniform(-1.5, 1.5, (K, 2))
    var = np.zeros((K, 2, 2))
    for k in range(K):
        a, d = rng.uniform(0.8, 12., 2)
        b = rng.uniform(-0.6, 0.6) * math.sqrt(a * d)
        var[k] = [[a, b], [b, d]]
    return amp, mean, var


CASES = [
    # (box, fx, fy, minval, start, K, seed, minradius)
    ((-20, 21, -18, 19), 0.3, -0.2, 1e-4, (0, 0), 2, 1, 3),
    ((-20, 21, -18, 19), 0.3, -0.2, 0., (0, 0), 3, 2, 3),
    ((-40, 33, -25, 47), -0.45, 0.1, 1e-6, (0, 1), 6, 3, 3),
    ((5, 30, -10, 12), 0., 0., 1e-3, (0, 0), 1, 4, 3),         # start left of box
    ((-30, -4, -20, 20), 0.5, 0.5, 1e-3, (0, 0), 2, 5, 1),     # start right of box
    ((-12, 12, 4, 28), 0.2, 0.1, 1e-3, (0, 0), 3, 6, 0),       # start below box
    ((-15, 16, -15, 16), 0., 0., 10., (0, 0), 1, 7, 2),        # minval above peak
    ((-3, 2, -2, 3), 0.1, 0.2, 1e-5, (0, 0), 2, 8, 3),         # tiny box
]


@pytest.mark.parametrize('case', range(len(CASES)))
def test_approx3_matches_oracle(case):
    from oracle import tractor_oracle as O
    from tractor_amd import gpu_render as G
    (x0, x1, y0, y1), fx, fy, minval, (xc, yc), K, seed, mr = CASES[case]
    amp, mean, var = _mix(K, seed)
    print('approx3 case', case, flush=True)
    r, xd, yd, bb = G.mog_render_approx3(amp, mean, var, x0, x1, y0, y1, fx, fy, minval, xc, yc,
                                         minradius=mr, derivs=True)
    rr, rxd, ryd, rbb = O.mog_grid_approx3(x0, x1, y0, y1, fx, fy, minval, amp, mean, var, xc, yc,
                                           minradius=mr)
    assert bb == rbb
    assert np.array_equal(r != 0, rr != 0)
    np.testing.assert_allclose(r, rr, rtol=1e-12, atol=1e-300)
    scale = max(np.abs(rxd).max(), np.abs(ryd).max(), 1e-300)
    np.testing.assert_allclose(xd, rxd, rtol=1e-11, atol=1e-13 * scale)
    np.testing.assert_allclose(yd, ryd, rtol=1e-11, atol=1e-13 * scale)


def test_mog_psf_minval_patch():
    """GaussianMixturePSF.getPointSourcePatch with minval > 0
    (tractor/psf.py:518-520 -> evaluate_grid_approx3): sliced patch whose
    pixels equal the oracle's and whose sum is within the truncation."""
    import tractor_amd as T
    from oracle import tractor_oracle as O
    psf = T.GaussianMixturePSF([0.8, 0.2], [[0., 0.], [0.3, -0.2]],
                               [[[4., 0.5], [0.5, 3.]], [[16., -2.], [-2., 12.]]])
    p = psf.getPointSourcePatch(40.3, 30.6, minval=1e-5)
    mog = psf.getMixtureOfGaussians()
    rr = 25
    x0, x1 = int(np.floor(40.3 - rr)), int(np.ceil(40.3 + rr)) + 1
    y0, y1 = int(np.floor(30.6 - rr)), int(np.ceil(30.6 + rr)) + 1
    cx, cy = int(mog.mean[0, 0] + 40.3), int(mog.mean[0, 1] + 30.6)
    res, _, _, (sx0, sx1, sy0, sy1) = O.mog_grid_approx3(x0, x1, y0, y1, 40.3, 30.6, 1e-5, mog.amp,
                                                         mog.mean, mog.var, cx, cy)
    assert (p.x0, p.y0) == (x0 + sx0, y0 + sy0)
    np.testing.assert_allclose(p.patch, res[sy0:sy1, sx0:sx1], rtol=1e-12, atol=0)
    assert abs(p.patch.sum() - 1.) < 1e-3


def test_galaxy_mog_path_with_minval():
    """A galaxy under a GaussianMixturePSF without a modelMask and with
    img.modelMinval > 0: run_mog -> mixture_to_patch -> approx3 ring growth
    (tractor/galaxy.py:307-333, mixture_profiles.py:425-438) on the no-mask
    box (halfsize rule, galaxy.py:273-290, 678-691), minval = modelMinval /
    counts; the sliced patch equals the oracle's approx3 of the convolved
    mixture and the model is counts times it."""
    import tractor_amd as T
    from oracle import tractor_oracle as O
    psf = T.GaussianMixturePSF([0.8, 0.2], [[0., 0.], [0.3, -0.2]],
                               [[[4., 0.5], [0.5, 3.]], [[16., -2.], [-2., 12.]]])
    H = W = 96
    tim = T.Image(data=np.zeros((H, W), np.float32), inverr=np.ones((H, W), np.float32), psf=psf)
    tim.modelMinval = 0.05
    px, py, flux, logre, ee1, ee2 = 47.3, 45.8, 300., np.log(3.), 0.2, -0.1
    gal = T.ExpGalaxy(T.PixPos(px, py), T.Flux(flux), T.EllipseESoft(logre, ee1, ee2))
    p = gal.getModelPatch(tim)
    print('minval galaxy patch', p.x0, p.y0, p.patch.shape, flush=True)
    # oracle: the sheared profile convolved with the PSF mixture, on the box
    amp, var = O.exp_mixture()
    vp = O.shear_mixture(var, np.eye(2) * 3600., O.ellipse_esoft_basis(logre, ee1, ee2))
    mog = psf.getMixtureOfGaussians()
    mean = np.zeros((len(amp), 2)) + np.array([px, py])
    camp, cmean, cvar = O.mog_convolve(amp, mean, vp, mog.amp, mog.mean, mog.var)
    hs = O.galaxy_patch_halfsize(4. * 3., px, py, (H, W), psf.getRadius())
    ox, _ = O.get_overlapping_region(int(np.floor(px - hs)), int(np.ceil(px + hs + 1)), 0, W - 1)
    oy, _ = O.get_overlapping_region(int(np.floor(py - hs)), int(np.ceil(py + hs + 1)), 0, H - 1)
    x0, x1, y0, y1 = ox.start, ox.stop, oy.start, oy.stop
    minval = 0.05 / flux
    res, _, _, (sx0, sx1, sy0, sy1) = O.mog_grid_approx3(
        x0, x1, y0, y1, 0., 0., minval, camp, cmean, cvar, int(cmean[0, 0]), int(cmean[0, 1]))
    ref = res[sy0:sy1, sx0:sx1]
    assert (p.x0, p.y0) == (x0 + sx0, y0 + sy0)
    assert p.patch.shape == ref.shape
    # FP64 approx3, the model cast to float32 after the counts product
    np.testing.assert_allclose(p.patch, (ref * flux).astype(np.float32), rtol=1e-6,
                               atol=1e-6 * flux * ref.max())
    # ring growth stops below 0.05 counts per pixel: the faint wings (about
    # 1.6% of the flux here) are cut, as in the reference
    assert 0.95 < p.patch.sum() / flux < 1.
