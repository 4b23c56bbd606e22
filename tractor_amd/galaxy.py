"""Exp / Dev / Gaussian galaxies (tractor/galaxy.py): MoG profiles rendered
through the FFT path (pixelized PSF) or the pixel-space MoG path (MoG PSF)
on the device, finite-difference derivatives as the reference takes them."""
import numpy as np

from . import mixture_profiles as mp
from .utils import ParamList, MultiParams
from .patch import Patch, ModelMask
from .pointsource import SingleProfileSource
from .render import get_overlapping_region
from . import gpu_render as G


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

    def getParamDerivatives(self, img, modelMask=None, get_patch0=False, **kwargs):
        """tractor/galaxy.py:137-223."""
        pos0 = self.getPosition()
        wcs = img.getWcs()
        px0, py0 = wcs.positionToPixel(pos0, self)
        counts = img.getPhotoCal().brightnessToCounts(self.brightness)
        minsb = img.modelMinval
        minval = minsb / counts if counts > 0 else None
        patch0 = self.getUnitFluxModelPatch(img, px=px0, py=py0, minval=minval,
                                            modelMask=modelMask, **kwargs)
        derivs = []
        if get_patch0:
            derivs.append(patch0)
        if patch0 is None:
            derivs.extend([None] * self.numberOfParams())
            return derivs
        if modelMask is None:
            x0, x1, y0, y1 = patch0.getExtent()
            modelMask = ModelMask.fromExtent(x0, x1, y0, y1)
        if self.isParamThawed('pos') and self.pos.numberOfParams() > 0:
            if counts == 0:
                derivs.extend([None] * len(pos0.getParams()))
            else:
                p0 = patch0.patch
                dx = np.zeros_like(p0)
                dx[1:-1, 1:-1] = (p0[1:-1, :-2] - p0[1:-1, 2:]) / 2.
                dy = np.zeros_like(p0)
                dy[1:-1, 1:-1] = (p0[:-2, 1:-1] - p0[2:, 1:-1]) / 2.
                derivs.extend(wcs.pixelDerivsToPositionDerivs(
                    pos0, self, counts, patch0, Patch(patch0.x0, patch0.y0, dx),
                    Patch(patch0.x0, patch0.y0, dy)))
        if not self.isParamFrozen('brightness'):
            bsteps = self.brightness.getStepSizes()
            params = self.brightness.getParams()
            for i, bstep in enumerate(bsteps):
                old = self.brightness.setParam(i, params[i] + bstep)
                ci = img.getPhotoCal().brightnessToCounts(self.brightness)
                self.brightness.setParam(i, old)
                if ci == counts:
                    df = None
                else:
                    df = patch0 * ((ci - counts) / bstep)
                    df.setName('d(%s)/d(bright%i)' % (self.dname, i))
                derivs.append(df)
        if not self.isParamFrozen('shape'):
            gsteps = self.shape.getStepSizes()
            gnames = self.shape.getParamNames()
            oldvals = self.shape.getParams()
            if counts == 0:
                derivs.extend([None] * len(oldvals))
                gsteps = []
            for i, gstep in enumerate(gsteps):
                old = self.shape.setParam(i, oldvals[i] + gstep)
                patchx = self.getUnitFluxModelPatch(img, px=px0, py=py0, minval=minval,
                                                    modelMask=modelMask, force_sourceOut=True,
                                                    **kwargs)
                self.shape.setParam(i, old)
                if patchx is None:
                    derivs.append(None)
                    continue
                dx = (patchx - patch0) * (counts / gstep)
                dx.setName('d(%s)/d(%s)' % (self.dname, gnames[i]))
                derivs.append(dx)
        return derivs


class ProfileGalaxy(object):
    """Renders a MoG profile (tractor/galaxy.py:225-519)."""

    def getUnitFluxModelPatch(self, img, px=None, py=None, minval=0.0, modelMask=None,
                              **kwargs):
        if px is None or py is None:
            px, py = img.getWcs().positionToPixel(self.getPosition(), self)
        patch = self._realGetUnitFluxModelPatch(img, px, py, minval, modelMask=modelMask,
                                                **kwargs)
        if patch is not None and modelMask is not None:
            assert patch.shape == modelMask.shape
        return patch

    def _realGetUnitFluxModelPatch(self, img, px, py, minval, modelMask=None,
                                   inner_real_nsigma=3., outer_real_nsigma=4.,
                                   force_halfsize=None, force_sourceOut=False, **kwargs):
        from .psf import HybridPSF
        if modelMask is not None:
            x0, y0 = modelMask.x0, modelMask.y0
        else:
            halfsize = self._getUnitFluxPatchSize(img, px=px, py=py, minval=minval)
            H, W = img.shape
            outx, inx = get_overlapping_region(int(np.floor(px - halfsize)),
                                               int(np.ceil(px + halfsize + 1)), 0, W - 1)
            outy, iny = get_overlapping_region(int(np.floor(py - halfsize)),
                                               int(np.ceil(py + halfsize + 1)), 0, H - 1)
            if inx == [] or iny == []:
                return None
            x0, x1 = outx.start, outx.stop
            y0, y1 = outy.start, outy.stop
        psf = img.getPsf()
        hybrid = isinstance(psf, HybridPSF)

        def run_mog(amix=None, mm=None):
            if amix is None:
                amix = self._getAffineProfile(img, px, py)
            if mm is None:
                mm = modelMask
            psfmix = psf.getMixtureOfGaussians(px=px, py=py)
            cmix = amix.convolve(psfmix)
            if mm is None:
                if minval is not None and minval > 0:
                    raise NotImplementedError('approximate (minval > 0) MoG rendering')
                return Patch(x0, y0, G.mog_render(cmix.amp, cmix.mean, cmix.var, x0, y0,
                                                  x1 - x0, y1 - y0))
            if mm.mask is not None:
                return Patch(mm.x0, mm.y0, G.mog_render(cmix.amp, cmix.mean, cmix.var, mm.x0,
                                                        mm.y0, mm.w, mm.h, mask=mm.mask,
                                                        masked_rules=True))
            return Patch(mm.x0, mm.y0, G.mog_render(cmix.amp, cmix.mean, cmix.var, mm.x0, mm.y0,
                                                    mm.w, mm.h))

        if hasattr(psf, 'getMixtureOfGaussians') and not hybrid:
            return run_mog(mm=modelMask)

        if modelMask is not None:
            mh, mw = modelMask.shape
            x1 = x0 + mw
            y1 = y0 + mh
            sourceOut = (px < x0 or px > x1 - 1 or py < y0 or py > y1 - 1)
            if sourceOut:
                neardx = neardy = 0.
                if px < x0:
                    neardx = x0 - px
                if px > x1:
                    neardx = px - x1
                if py < y0:
                    neardy = y0 - py
                if py > y1:
                    neardy = py - y1
                nearest = np.hypot(neardx, neardy)
                pixscale = img.wcs.pixscale_at((x0 + x1) / 2., (y0 + y1) / 2.)
                if nearest > (self.getRadius() / pixscale) and not force_sourceOut:
                    return None
                if hybrid:
                    return run_mog(mm=modelMask)
                raise RuntimeError('sourceOut, not hybrid')

        amix = self._getShearedProfile(img, px, py)
        # FFT size and the hybrid split use the PSF spectrum's size
        if modelMask is not None:
            from .render import mm_halfsize
            hs = mm_halfsize(px, py, (x0, y0, mw, mh), psf.shape)
            if force_halfsize is not None:
                hs = int(np.ceil(force_halfsize))
            nomask = None
            mmt = (x0, y0, mw, mh)
        else:
            hs = halfsize
            nomask = (x0, x1, y0, y1, halfsize)
            mmt = None
        pW = psf.getFourierTransformSize(hs)
        fftamp, fftvar = amix.amp, amix.var
        mogmix = None
        if hybrid and inner_real_nsigma is not None and outer_real_nsigma is not None:
            vv = amix.var[:, 0, 0] + amix.var[:, 1, 1]
            n1, n2 = inner_real_nsigma, outer_real_nsigma
            IM = ((pW / 2) ** 2 < (n2 ** 2 * vv))
            IF = ((pW / 2) ** 2 > (n1 ** 2 * vv))
            ramp = np.any(IM * IF)
            ns = (pW / 2) / np.maximum(1e-6, np.sqrt(vv))
            if np.any(IM):
                amps = amix.amp[IM]
                if ramp:
                    amps = amps * np.minimum(1., (n2 - ns[IM]) / (n2 - n1))
                mogmix = mp.MixtureOfGaussians(amps, amix.mean[IM, :] + np.array([px, py]),
                                               amix.var[IM], quick=True)
            if np.any(IF):
                amps = amix.amp[IF]
                if ramp:
                    amps = amps * np.minimum(1., (ns[IF] - n1) / (n2 - n1))
                fftamp, fftvar = amps, amix.var[IF]
            else:
                fftamp = None
        if fftamp is not None:
            pix = psf.pix if hybrid else psf
            if mmt is not None and force_halfsize is not None:
                raise NotImplementedError('force_halfsize with a modelMask')
            ix0, iy0, Gp = G.fft_render(fftamp, fftvar, px, py, pix, mm=mmt, nomask_extent=nomask)
        else:
            # no FFT components: an empty image placed as the FFT path would be
            from .render import galaxy_fft_job
            _, _, win = galaxy_fft_job(np.zeros(1), np.zeros((1, 2, 2)), px, py, psf.shape,
                                       mm=mmt, nomask_extent=nomask)
            ix0, iy0 = win[0], win[1]
            Gp = np.zeros((win[3], win[2]), np.float32)
        if mogmix is not None:
            gh, gw = Gp.shape
            mpatch = run_mog(amix=mogmix, mm=modelMask if modelMask is not None
                             else ModelMask(ix0, iy0, gw, gh))
            Gp = Gp + mpatch.patch
        return Patch(ix0, iy0, Gp)


class HoggGalaxy(ProfileGalaxy, Galaxy):
    """Affine / sheared profiles and the patch-size rule (galaxy.py:640-709)."""

    def getName(self):
        return 'HoggGalaxy'

    def getRadius(self):
        return self.nre * self.shape.re

    def _getAffineProfile(self, img, px, py):
        cdinv = img.getWcs().cdInverseAtPixel(px, py)
        Tinv = np.dot(cdinv, self.shape.getRaDecBasis())
        return self.getProfile().apply_affine(np.array([px, py]), Tinv)

    def _getShearedProfile(self, img, px, py):
        cdinv = img.getWcs().cdInverseAtPixel(px, py)
        Tinv = np.dot(cdinv, self.shape.getRaDecBasis())
        return self.getProfile().apply_shear(Tinv)

    def _getUnitFluxPatchSize(self, img, px=0., py=0., minval=0.):
        if hasattr(self, 'halfsize'):
            return self.halfsize
        pixscale = img.wcs.pixscale_at(px, py)
        halfsize = max(1., self.getRadius() / pixscale)
        h, w = img.shape
        imsize = max(px, w - 1 - px, py, h - 1 - py) + 1
        halfsize = min(halfsize, imsize)
        halfsize = max(halfsize, img.psf.getRadius())
        return int(np.ceil(halfsize))

    def getDerivativeShearedProfiles(self, img, px, py):
        derivs = []
        if self.isParamThawed('shape'):
            gsteps = self.shape.getStepSizes()
            gnames = self.shape.getParamNames()
            oldvals = self.shape.getParams()
            for i, gstep in enumerate(gsteps):
                old = self.shape.setParam(i, oldvals[i] + gstep)
                pro = self._getShearedProfile(img, px, py)
                self.shape.setParam(i, old)
                derivs.append(('shape.' + gnames[i], pro, gstep))
        return derivs


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
