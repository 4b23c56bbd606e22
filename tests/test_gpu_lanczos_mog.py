"""GPU parity of the standalone Lanczos export and of the pixel-space MoG
derivative outputs against the oracle.

- trh_lanczos_shift_batch (lanczos_shift_image, tractor/psf.py:26-58 ->
  lanczos_shift_3f, tractor/mp_fourier.i:405-429): float32 images, batched.
  Both sides run the same FP32 7-tap passes in the same tap order; the
  device fuses multiply and add, so the tolerance is 1e-6 of the image peak.
- trh_render_mog_batch with c_gauss_2d_masked rules and derivatives
  (tractor/gauss_masked.c, eval_all_dxy_f tractor/mix.i:85-142): values and
  d/dmean_x, d/dmean_y within 1e-5 of their peaks (FP32 on both sides; the
  device's exp differs from numpy's by an ulp or two) near the origin, and
  against the FP64 grid far from it (see that test).
"""
import math
import sys

import numpy as np
import pytest
import torch

import oracle as O
from tractor_amd import _lib
from tractor_amd import gpu_render as G
from tractor_amd._lib import check, ptr

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_err(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / max(np.max(np.abs(b)), 1e-30))


@pytest.mark.parametrize('H,W,n', [(5, 7, 3), (9, 9, 4), (33, 47, 5), (64, 64, 2), (63, 130, 3)])
def test_lanczos_shift_batch_vs_oracle(H, W, n):
    say('lanczos batch', H, W, n)
    assert torch.cuda.is_available()
    rng = np.random.default_rng(20261203 + H * 7 + W)
    imgs = rng.normal(size=(n, H, W)).astype(np.float32)
    imgs[0] = 0.
    imgs[0, H // 2, W // 2] = 1.                      # a delta: the taps themselves
    shifts = [(0., 0.), (0.5, -0.5), (-0.5, 0.25), (0.1234, 0.), (0., -0.377)]
    dx = np.array([shifts[i % len(shifts)][0] for i in range(n)])
    dy = np.array([shifts[i % len(shifts)][1] for i in range(n)])
    lib = _lib.load()
    h = _lib.handle(0)
    din = torch.from_numpy(imgs.ravel().copy()).cuda()
    dout = torch.full_like(din, float('nan'))
    ddx = torch.from_numpy(dx).cuda()
    ddy = torch.from_numpy(dy).cuda()
    check(lib.trh_lanczos_shift_batch(h.value, ptr(din), ptr(dout), n, H, W, ptr(ddx), ptr(ddy),
                                      None), 'trh_lanczos_shift_batch')
    torch.cuda.synchronize()
    got = dout.cpu().numpy().reshape(n, H, W)
    for i in range(n):
        ref = O.lanczos_shift_image(imgs[i], dx[i], dy[i])
        err = rel_err(got[i], ref)
        say('  image', i, 'shift', dx[i], dy[i], 'err', err)
        assert err <= 1e-6, (i, err)
        if dx[i] == 0. and dy[i] == 0.:
            assert np.array_equal(got[i], imgs[i])     # zero shift is the identity


def _psf_mog():
    # the 2-component PSF of the small configuration: 0.9 N(1.5) + 0.1 N(4.0)
    amp = np.array([0.9, 0.1])
    mean = np.zeros((2, 2))
    var = np.array([np.eye(2) * 1.5 ** 2, np.eye(2) * 4.0 ** 2])
    return amp, mean, var


def _convolved(kind, px, py):
    gamp, gvar = O.exp_mixture() if kind == 1 else O.dev_mixture()
    basis = O.ellipse_esoft_basis(math.log(2.5), 0.2, -0.1)
    vp = O.shear_mixture(gvar, np.eye(2) * 3600., basis)
    pamp, pmean, pvar = _psf_mog()
    mean = np.zeros((len(gamp), 2)) + np.array([px, py])
    return O.mog_convolve(gamp, mean, vp, pamp, pmean, pvar)


@pytest.mark.parametrize('kind', [1, 2])
def test_mog_masked_derivatives_vs_oracle(kind):
    """Near the image origin, where float32 pixel coordinates are exact to
    ~1e-6 px, the FP32 masked evaluator of the oracle (the reference's
    arithmetic) is the bar: values and derivatives within 1e-5 of peak."""
    say('mog masked derivatives, kind', kind)
    assert torch.cuda.is_available()
    rng = np.random.default_rng(20261203 + kind)
    camp, cmean, cvar = _convolved(kind, 23.37, 21.61)
    x0, y0, w, h = 0, 1, 47, 41
    mask = rng.uniform(size=(h, w)) > 0.2
    ref, rdx, rdy = O.mog_masked(x0, y0, mask, 0., 0., camp, cmean, cvar, derivs=True)
    got, gdx, gdy = G.mog_render(camp, cmean, cvar, x0, y0, w, h, mask=mask, derivs=True,
                                 masked_rules=True)
    e0, ex, ey = rel_err(got, ref), rel_err(gdx, rdx), rel_err(gdy, rdy)
    say('  value', e0, 'dx', ex, 'dy', ey)
    assert e0 <= 1e-5 and ex <= 1e-5 and ey <= 1e-5
    # masked pixels are zero in all three outputs
    assert not np.any(got[~mask]) and not np.any(gdx[~mask]) and not np.any(gdy[~mask])


@pytest.mark.parametrize('kind', [1, 2])
def test_mog_masked_derivatives_far_from_origin(kind):
    """At x ~ 1000 float32 coordinates carry ~6e-5 px of rounding, which the
    reference's FP32 evaluator (and the oracle's restatement of it) inherits;
    the kernel keeps each centre as an integer plus a fraction, so it is held
    to the FP64 grid evaluator instead: values within 1e-5 of peak, and the
    derivatives with respect to the mean within 1e-4 of peak of a central
    difference of that FP64 grid (moving every component by +d in x moves
    the image by +d: d/dmean_x = -d/dx of the image)."""
    say('mog masked derivatives far from the origin, kind', kind)
    rng = np.random.default_rng(20261213 + kind)
    camp, cmean, cvar = _convolved(kind, 1003.37, 2011.61)
    x0, y0, w, h = 980, 1990, 47, 41
    mask = rng.uniform(size=(h, w)) > 0.2
    got, gdx, gdy = G.mog_render(camp, cmean, cvar, x0, y0, w, h, mask=mask, derivs=True,
                                 masked_rules=True)
    ref = O.mog_grid(x0, x0 + w, y0, y0 + h, 0., 0., camp, cmean, cvar) * mask
    d = 1e-3
    fd = []
    for fx, fy in [(d, 0.), (0., d)]:
        fp = O.mog_grid(x0, x0 + w, y0, y0 + h, fx, fy, camp, cmean, cvar)
        fm = O.mog_grid(x0, x0 + w, y0, y0 + h, -fx, -fy, camp, cmean, cvar)
        fd.append(((fp - fm) / (2 * d)) * mask)
    e0, ex, ey = rel_err(got, ref), rel_err(gdx, fd[0]), rel_err(gdy, fd[1])
    say('  value', e0, 'dx', ex, 'dy', ey)
    assert e0 <= 1e-5 and ex <= 1e-4 and ey <= 1e-4


def test_mog_grid_rules_vs_oracle():
    """c_gauss_2d_grid rules (FP64 constants, skip d >= 100) on a window
    without a pixel map, the hybrid split's MoG part."""
    say('mog grid rules')
    gamp, gvar = O.dev_mixture()
    vp = O.shear_mixture(gvar, np.eye(2) * 3600., O.ellipse_esoft_basis(math.log(6.), -0.3, 0.2))
    pamp, pmean, pvar = _psf_mog()
    mean = np.zeros((len(gamp), 2)) + np.array([40.2, 37.9])
    camp, cmean, cvar = O.mog_convolve(gamp, mean, vp, pamp, pmean, pvar)
    ref = O.mog_grid(0, 80, 0, 76, 0., 0., camp, cmean, cvar)
    got = G.mog_render(camp, cmean, cvar, 0, 0, 80, 76)
    err = rel_err(got, ref)
    say('  err', err)
    assert err <= 1e-5
