"""Point sources (tractor/pointsource.py:9-241), rendered through
RenderPlans (tractor_amd.plan) like the galaxies."""
import numpy as np

from .plan import Pending, ready, resolve
from .utils import MultiParams


class BasicSource(object):
    def getPosition(self):
        return self.pos

    def setPosition(self, position):
        self.pos = position


class SingleProfileSource(BasicSource):
    """A source whose model is counts times one unit-flux patch
    (tractor/pointsource.py:17-60)."""

    def getBrightness(self):
        return self.brightness

    def setBrightness(self, brightness):
        self.brightness = brightness

    def getBrightnesses(self):
        return [self.getBrightness()]

    def getUnitFluxModelPatches(self, *args, **kwargs):
        return [self.getUnitFluxModelPatch(*args, **kwargs)]

    def _plan_unit_patches(self, plan, img, **kwargs):
        return [self._plan_unit(plan, img, **kwargs)]

    def getModelPatch(self, img, minsb=None, modelMask=None, **kwargs):
        return resolve(self._plan_model, img, minsb=minsb, modelMask=modelMask, **kwargs)

    def _plan_model(self, plan, img, minsb=None, modelMask=None, **kwargs):
        """counts x the unit-flux patch, with minval = minsb / counts
        (pointsource.py:36-60); None for zero or non-finite counts."""
        counts = img.getPhotoCal().brightnessToCounts(self.brightness)
        if counts == 0 or not np.isfinite(np.float32(counts)):
            return ready(None)
        if minsb is None:
            minsb = img.modelMinval
        unit = self._plan_unit(plan, img, minval=minsb / counts, modelMask=modelMask, **kwargs)
        if unit.extent is None:
            return ready(None)

        def done():
            u = unit.get()
            if u is None:
                return None
            if u.patch is not None:
                assert np.all(np.isfinite(u.patch))
            return u * counts
        return Pending(unit.extent, done)

    def _brightness_columns(self, img, counts):
        """Per thawed brightness parameter: None when stepping it leaves the
        counts in this image unchanged, else ('scale', dcounts / step, name)
        -- the unit patch times that factor (pointsource.py:221-237,
        galaxy.py:182-194)."""
        if self.isParamFrozen('brightness'):
            return []
        photo = img.getPhotoCal()
        vals = self.brightness.getParams()
        out = []
        for i, step in enumerate(self.brightness.getStepSizes(img)):
            old = self.brightness.setParam(i, vals[i] + step)
            ci = photo.brightnessToCounts(self.brightness)
            self.brightness.setParam(i, old)
            out.append(None if ci == counts else
                       ('scale', (ci - counts) / step,
                        'd(%s)/d(bright%i)' % (getattr(self, 'dname', 'ptsrc'), i)))
        return out


class PointSource(MultiParams, SingleProfileSource):
    """PointSource(pos, brightness) (tractor/pointsource.py:63-241)."""

    def __init__(self, pos, br):
        super(PointSource, self).__init__(pos, br)
        self.fixedRadius = None
        self.minRadius = None

    @staticmethod
    def getNamedParams():
        return dict(pos=0, brightness=1)

    def getSourceType(self):
        return 'PointSource'

    def __str__(self):
        return self.getSourceType() + ' at ' + str(self.pos) + ' with ' + str(self.brightness)

    def __repr__(self):
        return self.getSourceType() + '(' + repr(self.pos) + ', ' + repr(self.brightness) + ')'

    def copy(self):
        c = PointSource(self.pos.copy(), self.brightness.copy())
        c.liquid = list(self.liquid)
        c.fixedRadius = self.fixedRadius
        c.minRadius = self.minRadius
        return c

    def _getPsf(self, img):
        return img.getPsf()

    def getUnitFluxModelPatch(self, img, minval=0., derivs=False, modelMask=None, **kwargs):
        return resolve(self._plan_unit, img, minval=minval, derivs=derivs, modelMask=modelMask,
                       **kwargs)

    def _plan_unit(self, plan, img, minval=0., derivs=False, modelMask=None, **kwargs):
        """The PSF's point-source patch at the source (pointsource.py:
        102-125): None when the PSF's radius misses the image; clipped to the
        image when there is no modelMask."""
        from .psf import plan_point_patch
        px, py = img.getWcs().positionToPixel(self.getPosition(), self)
        H, W = img.shape
        psf = self._getPsf(img)
        r = psf.getRadius() if self.fixedRadius is None else self.fixedRadius
        if px + r < 0 or px - r > W or py + r < 0 or py - r > H:
            return ready(None)
        return plan_point_patch(psf, plan, px, py, minval=minval, radius=self.fixedRadius,
                                derivs=derivs, minradius=self.minRadius, modelMask=modelMask,
                                clipExtent=None if modelMask is not None else [0, W, 0, H])

    def getParamDerivatives(self, img, fastPosDerivs=True, modelMask=None, **kwargs):
        """Derivative patches of the thawed parameters (pointsource.py:
        130-241): position from the PSF's own (dx, dy) patches when it gives
        them, else forward differences of renders at the position steps;
        brightness by scaling the unit patch."""
        return resolve(self._plan_derivs, img, fastPosDerivs=fastPosDerivs,
                       modelMask=modelMask, **kwargs)

    def _plan_derivs(self, plan, img, fastPosDerivs=True, modelMask=None, **kwargs):
        n = self.numberOfParams()
        pos_thawed = not (self.isParamFrozen('pos') or self.pos.numberOfParams() == 0)
        photo = img.getPhotoCal()
        counts = photo.brightnessToCounts(self.brightness)
        bcols = self._brightness_columns(img, counts)
        if not pos_thawed and bcols and all(c is None for c in bcols):
            # nothing this image can constrain (pointsource.py:139-153)
            return ready([None] * n)
        minval = img.modelMinval / counts if counts > 0 else None
        want = pos_thawed and fastPosDerivs
        base = self._plan_unit(plan, img, minval=minval, derivs=want, modelMask=modelMask,
                               **kwargs)
        H, W = img.shape
        ext = base.extent
        if ext is None or ext[0] >= W or ext[0] + ext[2] <= 0 or ext[1] >= H or \
                ext[1] + ext[3] <= 0:
            return ready([None] * n)
        pos = self.getPosition()
        cols = []
        if pos_thawed:
            if want and base.has_derivs:
                cols.append(('own',))
            elif counts == 0:
                cols += [None] * pos.numberOfParams()
            else:
                vals = pos.getParams()
                for i, step in enumerate(pos.getStepSizes(img)):
                    old = pos.setParam(i, vals[i] + step)
                    r = self._plan_unit(plan, img, minval=minval, modelMask=modelMask, **kwargs)
                    pos.setParam(i, old)
                    cols.append(('fd', r, counts / step, 'd(ptsrc)/d(pos%i)' % i))
        cols += bcols
        wcs = img.getWcs()

        def done():
            got = base.get()
            p0 = got[0] if base.has_derivs else got
            out = []
            for c in cols:
                if c is None:
                    out.append(None)
                elif c[0] == 'own':
                    out.extend(wcs.pixelDerivsToPositionDerivs(pos, self, counts, p0, got[1],
                                                               got[2]))
                elif c[0] == 'scale':
                    d = p0 * c[1]
                    d.setName(c[2])
                    out.append(d)
                else:
                    q = c[1].get()
                    # a stepped render that left the image counts as zero
                    d = p0 * (-c[2]) if q is None else (q - p0) * c[2]
                    d.setName(c[3])
                    out.append(d)
            return out
        return Pending(ext, done)
