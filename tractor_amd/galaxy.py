"""Exp / Dev / Gaussian galaxies (tractor/galaxy.py).

Every render is planned into a RenderPlan (tractor_amd.plan): the
reference's extent, FFT-size and hybrid-split rules run on the host when the
render is planned, the pixels come from one batched device call per bucket.
The single-call methods (getUnitFluxModelPatch, getModelPatch,
getParamDerivatives) plan, run and read one plan; Tractor.getModelImage,
getDerivs and the optimizers' line searches plan every source together.
"""
import math

import numpy as np

from . import mixture_profiles as mp
from . import render as R
from .patch import Patch, ModelMask
from .plan import Pending, ready, resolve
from .pointsource import SingleProfileSource
from .render import get_overlapping_region
from .utils import ParamList, MultiParams


class GalaxyShape(ParamList):
    """(re [arcsec], ab, phi [deg]) (tractor/galaxy.py:37-88)."""

    @staticmethod
    def getName():
        return 'Galaxy Shape'

    @staticmethod
    def getNamedParams():
        return dict(re=0, ab=1, phi=2)

    def __repr__(self):
        return 're=%g, ab=%g, phi=%g' % (self.re, self.ab, self.phi)

    def __str__(self):
        return '%s: re=%.2f, ab=%.2f, phi=%.1f' % (self.getName(), self.re, self.ab, self.phi)

    def getRaDecBasis(self):
        phi = np.deg2rad(90 - self.phi)
        re_deg = max(1. / 30, self.re) / 3600.
        cp, sp = np.cos(phi), np.sin(phi)
        return re_deg * np.array([[cp, sp * self.ab], [-sp, cp * self.ab]])


def _image_box(img, px, py, halfsize):
    """The no-modelMask render box (tractor/galaxy.py:273-290): the square
    of half-size `halfsize` about (px, py), clipped to the image, as
    (x0, x1, y0, y1), or None when it misses the image."""
    H, W = img.shape
    xs, xin = get_overlapping_region(int(np.floor(px - halfsize)),
                                     int(np.ceil(px + halfsize + 1)), 0, W - 1)
    ys, yin = get_overlapping_region(int(np.floor(py - halfsize)),
                                     int(np.ceil(py + halfsize + 1)), 0, H - 1)
    if xin == [] or yin == []:
        return None
    return xs.start, xs.stop, ys.start, ys.stop


def _distance_outside(px, py, mm):
    """How far (px, py) lies outside the modelMask, or None when it is
    inside (the source-out test of tractor/galaxy.py:354-365)."""
    x0, y0, x1, y1 = mm.x0, mm.y0, mm.x0 + mm.w, mm.y0 + mm.h
    if x0 <= px <= x1 - 1 and y0 <= py <= y1 - 1:
        return None
    ddx = (x0 - px) if px < x0 else ((px - x1) if px > x1 else 0.)
    ddy = (y0 - py) if py < y0 else ((py - y1) if py > y1 else 0.)
    return math.hypot(ddx, ddy)


def hybrid_split(amp, var, pW, nsigma_in, nsigma_out):
    """Split a sheared profile between the FFT and the pixel-space mixture
    for a HybridPSF (tractor/galaxy.py:412-455): components wider than
    nsigma_out sigma at the FFT box's edge (pW / 2) go to the mixture, those
    narrower than nsigma_in to the FFT, with linear weights in between.
    Returns ((fft_amp, fft_var) or None, (mog_amp, index) or None)."""
    vv = var[:, 0, 0] + var[:, 1, 1]
    half2 = (pW / 2.) ** 2
    to_mog = half2 < nsigma_out ** 2 * vv
    to_fft = half2 > nsigma_in ** 2 * vv
    both = np.any(to_mog & to_fft)
    ns = (pW / 2.) / np.maximum(1e-6, np.sqrt(vv))
    width = nsigma_out - nsigma_in
    mog = fft = None
    if np.any(to_mog):
        a = amp[to_mog]
        if both:
            a = a * np.minimum(1., (nsigma_out - ns[to_mog]) / width)
        mog = (a, to_mog)
    if np.any(to_fft):
        a = amp[to_fft]
        if both:
            a = a * np.minimum(1., (ns[to_fft] - nsigma_in) / width)
        fft = (a, var[to_fft])
    return fft, mog


class Galaxy(MultiParams, SingleProfileSource):
    """Position, brightness, shape (tractor/galaxy.py:91-223)."""

    def __init__(self, *args):
        super(Galaxy, self).__init__(*args)
        self.name = self.getName()
        self.dname = self.getDName()

    @staticmethod
    def getNamedParams():
        return dict(pos=0, brightness=1, shape=2)

    def getName(self):
        return 'Galaxy'

    def getDName(self):
        return 'gal'

    def getSourceType(self):
        return self.name

    def getShape(self):
        return self.shape

    def __str__(self):
        return (self.name + ' at ' + str(self.pos) + ' with ' + str(self.brightness) + ' and ' +
                str(self.shape))

    def __repr__(self):
        return (self.name + '(pos=' + repr(self.pos) + ', brightness=' + repr(self.brightness) +
                ', shape=' + repr(self.shape) + ')')

    def copy(self):
        c = self.__class__(self.pos.copy(), self.brightness.copy(), self.shape.copy())
        c.liquid = list(self.liquid)
        return c

    def getParamDerivatives(self, img, modelMask=None, **kwargs):
        """Derivative patches of the thawed parameters (tractor/galaxy.py:
        137-223): position by central differences of the unit render through
        the WCS, brightness by scaling it, shape by forward differences of
        renders on the unit render's window (force_sourceOut)."""
        return resolve(self._plan_derivs, img, modelMask=modelMask, **kwargs)

    def _stepped(self, plan, img, param, i, step, **kw):
        """A unit render planned with parameter i of `param` moved by step
        (the parameter is restored at once: the plan has read it)."""
        vals = param.getParams()
        old = param.setParam(i, vals[i] + step)
        try:
            return self._plan_unit(plan, img, **kw)
        finally:
            param.setParam(i, old)

    def _plan_derivs(self, plan, img, modelMask=None, **kwargs):
        wcs = img.getWcs()
        pos0 = self.getPosition()
        px, py = wcs.positionToPixel(pos0, self)
        counts = img.getPhotoCal().brightnessToCounts(self.brightness)
        minval = img.modelMinval / counts if counts > 0 else None
        base = self._plan_unit(plan, img, px=px, py=py, minval=minval, modelMask=modelMask,
                               **kwargs)
        n = self.numberOfParams()
        if base.extent is None:
            return ready([None] * n)
        window = modelMask if modelMask is not None else ModelMask(*base.extent)
        # one entry per thawed parameter, in parameter order: None, or how
        # to make the column from the unit render
        cols = []
        if self.isParamThawed('pos') and self.pos.numberOfParams() > 0:
            npos = len(pos0.getParams())
            cols += [None] * npos if counts == 0 else [('pos', npos)]
        cols += self._brightness_columns(img, counts)
        if not self.isParamFrozen('shape'):
            names = self.shape.getParamNames()
            if counts == 0:
                cols += [None] * len(names)
            else:
                for i, step in enumerate(self.shape.getStepSizes()):
                    r = self._stepped(plan, img, self.shape, i, step, px=px, py=py,
                                      minval=minval, modelMask=window, force_sourceOut=True,
                                      **kwargs)
                    cols.append(('diff', r, counts / step,
                                 'd(%s)/d(%s)' % (self.dname, names[i])))
        cols += self._extra_columns(plan, img, px, py, modelMask, counts, kwargs)

        def done():
            p0 = base.get()
            out = []
            for c in cols:
                if c is None:
                    out.append(None)
                elif c[0] == 'pos':
                    gx, gy = _central(p0)
                    out.extend(wcs.pixelDerivsToPositionDerivs(pos0, self, counts, p0, gx, gy))
                elif c[0] == 'scale':
                    d = p0 * c[1]
                    d.setName(c[2])
                    out.append(d)
                else:
                    q = c[1].get()
                    if q is None:
                        out.append(None)
                        continue
                    d = (q - p0) * c[2]
                    d.setName(c[3])
                    out.append(d)
            return out
        return Pending(base.extent, done)

    def _extra_columns(self, plan, img, px, py, modelMask, counts, kwargs):
        """Columns of parameters beyond position, brightness and shape (the
        Sersic index); none here."""
        return []


def _central(p0):
    from .psf import central_differences
    gx, gy = central_differences(p0.patch)
    return Patch(p0.x0, p0.y0, gx), Patch(p0.x0, p0.y0, gy)


class ProfileGalaxy(object):
    """Renders a MoG profile (tractor/galaxy.py:225-519)."""

    def getUnitFluxModelPatch(self, img, px=None, py=None, minval=0.0, modelMask=None,
                              **kwargs):
        patch = resolve(self._plan_unit, img, px=px, py=py, minval=minval,
                        modelMask=modelMask, **kwargs)
        if patch is not None and modelMask is not None:
            assert patch.shape == modelMask.shape
        return patch

    def _plan_mog(self, plan, img, psf, px, py, amix, mm, box, minval):
        """The profile `amix` (means already at the source) convolved with
        the PSF's mixture and evaluated in pixel space (run_mog,
        tractor/galaxy.py:307-334): on the modelMask (masked rules when it
        has a pixel map, else the grid), or without one on `box` =
        (x0, x1, y0, y1) by mixture_to_patch (grid, or ring growth when
        minval > 0)."""
        cmix = amix.convolve(psf.getMixtureOfGaussians(px=px, py=py))
        if mm is None:
            return mp.plan_mixture(plan, cmix, box[0], box[1], box[2], box[3], minval)
        masked = mm.mask is not None
        slot = plan.mog(cmix.amp, cmix.mean, cmix.var, mm.x0, mm.y0, mm.w, mm.h,
                        mask=mm.mask, masked_rules=masked)
        x0, y0 = mm.x0, mm.y0
        return Pending((x0, y0, mm.w, mm.h), lambda: Patch(x0, y0, plan.result(slot)))

    def _plan_unit(self, plan, img, px=None, py=None, minval=0., modelMask=None,
                   inner_real_nsigma=3., outer_real_nsigma=4., force_halfsize=None,
                   force_sourceOut=False, **kwargs):
        """Plan the unit-flux render of _realGetUnitFluxModelPatch
        (tractor/galaxy.py:266-519)."""
        from .psf import HybridPSF
        if px is None or py is None:
            px, py = img.getWcs().positionToPixel(self.getPosition(), self)
        psf = img.getPsf()
        hybrid = isinstance(psf, HybridPSF)
        mm = modelMask
        box = halfsize = None
        if mm is None:
            halfsize = self._getUnitFluxPatchSize(img, px=px, py=py, minval=minval)
            box = _image_box(img, px, py, halfsize)
            if box is None:
                return ready(None)
        # a mixture PSF: the analytic convolution
        if hasattr(psf, 'getMixtureOfGaussians') and not hybrid:
            return self._plan_mog(plan, img, psf, px, py, self._getAffineProfile(img, px, py),
                                  mm, box, minval)
        if mm is not None:
            far = _distance_outside(px, py, mm)
            if far is not None:
                pixscale = img.wcs.pixscale_at(mm.x0 + mm.w / 2., mm.y0 + mm.h / 2.)
                if far > self.getRadius() / pixscale and not force_sourceOut:
                    return ready(None)
                if not hybrid:
                    raise RuntimeError('sourceOut, not hybrid')
                return self._plan_mog(plan, img, psf, px, py,
                                      self._getAffineProfile(img, px, py), mm, None, minval)
            mmt = (mm.x0, mm.y0, mm.w, mm.h)
            nomask = None
            hs = (R.mm_halfsize(px, py, mmt, psf.shape) if force_halfsize is None
                  else int(np.ceil(force_halfsize)))
        else:
            mmt = None
            nomask = tuple(box) + (halfsize,)
            hs = halfsize
        sheared = self._getShearedProfile(img, px, py)
        fft = (sheared.amp, sheared.var)
        mog = None
        if hybrid and inner_real_nsigma is not None and outer_real_nsigma is not None:
            fft, mog = hybrid_split(sheared.amp, sheared.var, psf.getFourierTransformSize(hs),
                                    inner_real_nsigma, outer_real_nsigma)
        pix = psf.pix if hybrid else psf
        amp, var = fft if fft is not None else (np.zeros(1), np.zeros((1, 2, 2)))
        sz, job, (X0, Y0, w, h) = R.galaxy_fft_job(np.asarray(amp), np.asarray(var), px, py,
                                                   psf.shape, mm=mmt, nomask_extent=nomask,
                                                   halfsize=hs if mmt is not None else None)
        slot = plan.fft(pix, sz, job, w, h) if fft is not None else None
        part = None
        if mog is not None:
            a, idx = mog
            mmix = mp.MixtureOfGaussians(a, sheared.mean[idx] + np.array([px, py]),
                                         sheared.var[idx], quick=True)
            part = self._plan_mog(plan, img, psf, px, py, mmix,
                                  mm if mm is not None else ModelMask(X0, Y0, w, h), None, 0.)

        def done():
            G = plan.result(slot) if slot is not None else np.zeros((h, w), np.float32)
            if part is not None:
                G = G + part.get().patch
            return Patch(X0, Y0, G)
        return Pending((X0, Y0, w, h), done)


class HoggGalaxy(ProfileGalaxy, Galaxy):
    """Affine / sheared profiles and the patch-size rule (galaxy.py:640-709)."""

    def getName(self):
        return 'HoggGalaxy'

    def getRadius(self):
        return self.nre * self.shape.re

    def _basis_in_pixels(self, img, px, py):
        return np.dot(img.getWcs().cdInverseAtPixel(px, py), self.shape.getRaDecBasis())

    def _getAffineProfile(self, img, px, py):
        return self.getProfile().apply_affine(np.array([px, py]),
                                              self._basis_in_pixels(img, px, py))

    def _getShearedProfile(self, img, px, py):
        return self.getProfile().apply_shear(self._basis_in_pixels(img, px, py))

    def _getUnitFluxPatchSize(self, img, px=0., py=0., minval=0.):
        """The no-modelMask half-size (galaxy.py:678-691): nre r_e in
        pixels, no more than the distance to the far image edge, no less than
        the PSF radius."""
        if hasattr(self, 'halfsize'):
            return self.halfsize
        h, w = img.shape
        far_edge = max(px, w - 1 - px, py, h - 1 - py) + 1
        r = min(max(1., self.getRadius() / img.wcs.pixscale_at(px, py)), far_edge)
        return int(np.ceil(max(r, img.psf.getRadius())))

    def getDerivativeShearedProfiles(self, img, px, py):
        """(name, sheared profile at the stepped shape, step) per thawed
        shape parameter (tractor/galaxy.py:693-709)."""
        out = []
        if self.isParamThawed('shape'):
            names = self.shape.getParamNames()
            for i, step in enumerate(self.shape.getStepSizes()):
                vals = self.shape.getParams()
                old = self.shape.setParam(i, vals[i] + step)
                pro = self._getShearedProfile(img, px, py)
                self.shape.setParam(i, old)
                out.append(('shape.' + names[i], pro, step))
        return out


class GaussianGalaxy(HoggGalaxy):
    nre = 6.
    profile = mp.MixtureOfGaussians(np.array([1.]), np.zeros((1, 2)), np.array([[[1., 0.], [0., 1.]]]))

    def __init__(self, *args, **kwargs):
        self.nre = GaussianGalaxy.nre
        super(GaussianGalaxy, self).__init__(*args, **kwargs)

    def getName(self):
        return 'GaussianGalaxy'

    def getProfile(self):
        return GaussianGalaxy.profile


class ExpGalaxy(HoggGalaxy):
    nre = 4.
    profile = mp.get_exp_mixture()

    def __init__(self, *args, **kwargs):
        self.nre = ExpGalaxy.nre
        super(ExpGalaxy, self).__init__(*args, **kwargs)

    def getName(self):
        return 'ExpGalaxy'

    def getProfile(self):
        return ExpGalaxy.profile


class DevGalaxy(HoggGalaxy):
    nre = 8.
    profile = mp.get_dev_mixture()

    def __init__(self, *args, **kwargs):
        self.nre = DevGalaxy.nre
        super(DevGalaxy, self).__init__(*args, **kwargs)

    def getName(self):
        return 'DevGalaxy'

    def getProfile(self):
        return DevGalaxy.profile
