"""Point sources (tractor/pointsource.py:9-241)."""
import numpy as np

from .utils import MultiParams


class BasicSource(object):
    def getPosition(self):
        return self.pos

    def setPosition(self, position):
        self.pos = position


class SingleProfileSource(BasicSource):
    """getModelPatch = counts x unit-flux patch (tractor/pointsource.py:17-60)."""

    def getBrightness(self):
        return self.brightness

    def setBrightness(self, brightness):
        self.brightness = brightness

    def getBrightnesses(self):
        return [self.getBrightness()]

    def getUnitFluxModelPatches(self, *args, **kwargs):
        return [self.getUnitFluxModelPatch(*args, **kwargs)]

    def getModelPatch(self, img, minsb=None, modelMask=None, **kwargs):
        counts = img.getPhotoCal().brightnessToCounts(self.brightness)
        if counts == 0:
            return None
        if not np.isfinite(np.float32(counts)):
            return None
        if minsb is None:
            minsb = img.modelMinval
        minval = minsb / counts
        upatch = self.getUnitFluxModelPatch(img, minval=minval, modelMask=modelMask, **kwargs)
        if upatch is None:
            return None
        if upatch.patch is not None:
            assert np.all(np.isfinite(upatch.patch))
        return upatch * counts


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
        (px, py) = img.getWcs().positionToPixel(self.getPosition(), self)
        H, W = img.shape
        psf = self._getPsf(img)
        r = self.fixedRadius
        if r is None:
            r = psf.getRadius()
        if px + r < 0 or px - r > W or py + r < 0 or py - r > H:
            return None
        clipExtent = None
        if modelMask is None:
            clipExtent = [0, W, 0, H]
        return psf.getPointSourcePatch(px, py, minval=minval, radius=self.fixedRadius,
                                       derivs=derivs, minradius=self.minRadius,
                                       modelMask=modelMask, clipExtent=clipExtent)

    def getParamDerivatives(self, img, fastPosDerivs=True, modelMask=None, **kwargs):
        """tractor/pointsource.py:130-241."""
        pos_frozen = self.isParamFrozen('pos') or (self.pos.numberOfParams() == 0)
        bright_frozen = self.isParamFrozen('brightness')
        photo = img.getPhotoCal()
        counts0 = photo.brightnessToCounts(self.brightness)
        if pos_frozen and not bright_frozen:
            bsteps = self.brightness.getStepSizes(img)
            bvals = self.brightness.getParams()
            allzero = True
            for i, bstep in enumerate(bsteps):
                old = self.brightness.setParam(i, bvals[i] + bstep)
                ci = photo.brightnessToCounts(self.brightness)
                self.brightness.setParam(i, old)
                if ci != counts0:
                    allzero = False
                    break
            if allzero:
                return [None] * self.numberOfParams()
        pos = self.getPosition()
        wcs = img.getWcs()
        minsb = img.modelMinval
        minval = minsb / counts0 if counts0 > 0 else None
        derivs = (not pos_frozen) and fastPosDerivs
        patchdx = patchdy = None
        if derivs:
            patches = self.getUnitFluxModelPatch(img, minval=minval, derivs=True,
                                                 modelMask=modelMask, **kwargs)
            if patches is None:
                return [None] * self.numberOfParams()
            if not isinstance(patches, tuple):
                patch0 = patches
            else:
                patch0, patchdx, patchdy = patches
        else:
            patch0 = self.getUnitFluxModelPatch(img, minval=minval, modelMask=modelMask, **kwargs)
        if patch0 is None:
            return [None] * self.numberOfParams()
        H, W = img.shape
        if not patch0.overlapsBbox((0, W, 0, H)):
            return [None] * self.numberOfParams()
        out = []
        if not pos_frozen:
            if patchdx is not None and patchdy is not None:
                out.extend(wcs.pixelDerivsToPositionDerivs(pos, self, counts0, patch0, patchdx,
                                                           patchdy))
            elif counts0 == 0:
                out.extend([None] * pos.numberOfParams())
            else:
                psteps = pos.getStepSizes(img)
                pvals = pos.getParams()
                for i, pstep in enumerate(psteps):
                    old = pos.setParam(i, pvals[i] + pstep)
                    patchx = self.getUnitFluxModelPatch(img, minval=minval, modelMask=modelMask,
                                                        **kwargs)
                    pos.setParam(i, old)
                    if patchx is None:
                        dx = patch0 * (-1 * counts0 / pstep)
                    else:
                        dx = (patchx - patch0) * (counts0 / pstep)
                    dx.setName('d(ptsrc)/d(pos%i)' % i)
                    out.append(dx)
        if not bright_frozen:
            bsteps = self.brightness.getStepSizes(img)
            bvals = self.brightness.getParams()
            for i, bstep in enumerate(bsteps):
                old = self.brightness.setParam(i, bvals[i] + bstep)
                ci = photo.brightnessToCounts(self.brightness)
                self.brightness.setParam(i, old)
                if ci == counts0:
                    df = None
                else:
                    df = patch0 * ((ci - counts0) / bstep)
                    df.setName('d(ptsrc)/d(bright%i)' % i)
                out.append(df)
        return out
