"""Mixture-of-Gaussians galaxy profiles and affine/shear transforms.

The profile tables are the reference's (tractor/mixture_profiles.py:17-38):
renders must match it exactly, so the numbers are the same.
"""
import math

import numpy as np

_exp_amp = np.array([2.34853813e-03, 3.07995260e-02, 2.23364214e-01,
                     1.17949102e+00, 4.33873750e+00, 5.99820770e+00])
_exp_var = np.array([1.20078965e-03, 8.84526493e-03, 3.91463084e-02,
                     1.39976817e-01, 4.60962500e-01, 1.50159566e+00])
_dev_amp = np.array([4.26347652e-02, 2.40127183e-01, 6.85907632e-01,
                     1.51937350e+00, 2.83627243e+00, 4.46467501e+00,
                     5.72440830e+00, 5.60989349e+00])
_dev_var = np.array([2.23759216e-04, 1.00220099e-03, 4.18731126e-03,
                     1.69432589e-02, 6.84850479e-02, 2.87207080e-01,
                     1.33320254e+00, 8.40215071e+00])
_dev_core = 0.010233


class MixtureOfGaussians(object):
    """amp (K,), mean (K,2), var (K,2,2) (tractor/mixture_profiles.py:47-66)."""
    def __init__(self, amp, mean, var, quick=False):
        if quick:
            self.amp, self.mean, self.var = amp, mean, var
        else:
            self.amp = np.atleast_1d(amp).astype(float)
            self.mean = np.atleast_2d(np.array(mean)).astype(float)
            var = np.asarray(var, float)
            K = self.amp.shape[0]
            if var.size == K:
                v = np.zeros((K, 2, 2))
                v[:, 0, 0] = var
                v[:, 1, 1] = var
                var = v
            var = var.copy()
            off = 0.5 * (var[:, 0, 1] + var[:, 1, 0])
            var[:, 0, 1] = var[:, 1, 0] = off
            self.var = var
        self.K, self.D = self.mean.shape

    def copy(self):
        return MixtureOfGaussians(self.amp.copy(), self.mean.copy(), self.var.copy(), quick=True)

    def normalize(self):
        self.amp = self.amp / np.sum(self.amp)

    def apply_affine(self, shift, scale):
        """mixture_profiles.py:128-144: mean + shift, scale var scale^T."""
        return MixtureOfGaussians(self.amp, self.mean + shift, scale @ self.var @ scale.T,
                                  quick=True)

    def apply_shear(self, scale):
        """mixture_profiles.py:146-155."""
        return MixtureOfGaussians(self.amp, self.mean, scale @ self.var @ scale.T, quick=True)

    def convolve(self, other):
        """mixture_profiles.py:158-172 (other's components are the outer loop)."""
        K, Ko = self.K, other.K
        amp = np.zeros(K * Ko)
        mean = np.zeros((K * Ko, 2))
        var = np.zeros((K * Ko, 2, 2))
        for k in range(Ko):
            sl = slice(k * K, (k + 1) * K)
            amp[sl] = self.amp * other.amp[k]
            mean[sl] = self.mean + other.mean[k]
            var[sl] = self.var + other.var[k]
        return MixtureOfGaussians(amp, mean, var, quick=True)

    def evaluate_grid(self, x0, x1, y0, y1, cx, cy):
        """Exact grid (evaluate_grid_dstn, tractor/mixture_profiles.py:270-283,
        c_gauss_2d_grid) on the device -> Patch."""
        from .plan import resolve
        return resolve(self.plan_grid, x0, x1, y0, y1, cx, cy)

    def plan_grid(self, plan, x0, x1, y0, y1, cx, cy):
        """evaluate_grid as a Pending in `plan`."""
        from .patch import Patch
        from .plan import Pending
        slot = plan.mog(self.amp, self.mean, self.var, x0, y0, x1 - x0, y1 - y0, cx, cy)
        return Pending((x0, y0, x1 - x0, y1 - y0), lambda: Patch(x0, y0, plan.result(slot)))

    def evaluate_grid_approx3(self, x0, x1, y0, y1, fx, fy, minval, derivs=False, minradius=3,
                              doslice=True, maxmargin=100):
        """Ring-growth evaluation that stops below minval
        (tractor/mixture_profiles.py:345-411 -> c_gauss_2d_approx3) on the
        device.  Returns a Patch (or (patch, dx, dy) with derivs), sliced to
        the evaluated pixels when doslice; None when the first component lies
        more than maxmargin pixels outside the box."""
        from .patch import Patch
        from . import gpu_render as G
        cx = int(self.mean[0, 0] + fx)
        cy = int(self.mean[0, 1] + fy)
        if (cx < x0 - maxmargin or cx > x1 + maxmargin or
                cy < y0 - maxmargin or cy > y1 + maxmargin):
            return None
        res, xd, yd, (sx0, sx1, sy0, sy1) = G.mog_render_approx3(
            self.amp, self.mean, self.var, int(x0), int(x1), int(y0), int(y1), float(fx), float(fy),
            float(minval), cx, cy, int(minradius), derivs=derivs)
        if doslice:
            slc = slice(sy0, sy1), slice(sx0, sx1)
            res = res[slc].copy()
            if derivs:
                xd = xd[slc].copy()
                yd = yd[slc].copy()
            x0 += sx0
            y0 += sy0
        if derivs:
            return Patch(x0, y0, res), Patch(x0, y0, xd), Patch(x0, y0, yd)
        return Patch(x0, y0, res)


def mixture_to_patch(mixture, x0, x1, y0, y1, minval=0., exactExtent=False):
    """tractor/mixture_profiles.py:425-438."""
    from .plan import resolve
    return resolve(plan_mixture, mixture, x0, x1, y0, y1, minval, exactExtent)


def plan_mixture(plan, mixture, x0, x1, y0, y1, minval=0., exactExtent=False):
    """mixture_to_patch as a Pending in `plan`: the exact grid at minval 0,
    else the ring growth of approx3 (evaluated at once: its window is the
    evaluated region)."""
    from .plan import ready
    if minval == 0. or minval is None:
        return mixture.plan_grid(plan, x0, x1, y0, y1, 0., 0.)
    p = mixture.evaluate_grid_approx3(x0, x1, y0, y1, 0., 0., minval, doslice=not exactExtent)
    if p is None:
        return ready(None)
    return ready(p, (p.x0, p.y0, p.patch.shape[1], p.patch.shape[0]))


def get_exp_mixture():
    amp = _exp_amp / np.sum(_exp_amp)
    m = MixtureOfGaussians(amp, np.zeros((6, 2)), _exp_var)
    m.normalize()
    return m


def get_dev_mixture():
    amp = np.append(_dev_core, _dev_amp * (1. - _dev_core) / np.sum(_dev_amp))
    m = MixtureOfGaussians(amp, np.zeros((9, 2)), np.append(0., _dev_var))
    m.normalize()
    return m


def esoft_basis(logre, ee1, ee2):
    """EllipseESoft -> EllipseE.getRaDecBasis (tractor/ellipses.py:126-155,
    232-255): r_e unit vectors -> degrees."""
    re = math.exp(logre)
    ee = math.hypot(ee1, ee2)
    e = 1. - math.exp(-ee)
    return e_basis(re, e, math.atan2(ee2, ee1) / 2.)


def e_basis(re, e, theta):
    ct, st = math.cos(theta), math.sin(theta)
    maxab = 1000.
    ab = maxab if e >= 1. else min(maxab, (1. + e) / (1. - e))
    r_deg = re / 3600.
    return r_deg * np.array([[ct / ab, st], [-st / ab, ct]])
