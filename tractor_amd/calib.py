"""Calibration and position objects: NullWCS, PixPos, Flux / Fluxes /
NanoMaggies, NullPhotoCal / LinearPhotoCal, NullSky / ConstantSky
(tractor/wcs.py:8-49, 373-390; tractor/brightness.py; tractor/sky.py;
tractor/ducks.py:347-438)."""
import numpy as np

from .utils import BaseParams, ParamList, ScalarParam, MultiParams, getClassName
from .patch import Patch


class WCS(object):
    """Duck-type base (tractor/ducks.py:347-438)."""

    def pixelDerivsToPositionDerivs(self, pos, src, counts0, patch0, patchdx, patchdy):
        """(dx * cdi[0,i] + dy * cdi[1,i]) * counts per thawed position
        parameter (tractor/ducks.py:413-423)."""
        cdi = self.cdInverseAtPosition(pos, src=src)
        derivs = []
        for i, pname in pos.getThawedParamIndicesAndNames():
            d = (patchdx * cdi[0, i] + patchdy * cdi[1, i]) * counts0
            d.setName('d(ptsrc)/d(pos.%s)' % pname)
            derivs.append(d)
        return derivs

    def cdInverseAtPosition(self, pos, src=None):
        px, py = self.positionToPixel(pos, src=src)
        return self.cdInverseAtPixel(px, py)

    def pixscale_at(self, x, y):
        return 3600. * np.sqrt(np.abs(np.linalg.det(self.cdAtPixel(x, y))))


class NullWCS(BaseParams, WCS):
    """Identity WCS with a nominal pixel scale (tractor/wcs.py:8-49).

    cd = I * pixscale/3600, so position derivatives carry cd^-1 =
    3600/pixscale * I, exactly as the reference's NullWCS does."""

    def __init__(self, pixscale=1., dx=0., dy=0.):
        self.dx = dx
        self.dy = dy
        self.setPixscale(pixscale)

    def hashkey(self):
        return ('NullWCS', self.dx, self.dy)

    def copy(self):
        return NullWCS(self.pixscale, self.dx, self.dy)

    def setPixscale(self, pixscale):
        self.pixscale = pixscale
        self.cd = np.array([[1., 0.], [0., 1.]]) * self.pixscale / 3600.
        self.cd_inverse = np.array([[1., 0.], [0., 1.]]) / (self.pixscale / 3600.)

    def positionToPixel(self, pos, src=None):
        return pos.x + self.dx, pos.y + self.dy

    def pixelToPosition(self, x, y, src=None):
        return x - self.dx, y - self.dy

    def cdAtPixel(self, x, y):
        return self.cd

    def cdInverseAtPixel(self, x, y):
        return self.cd_inverse

    def cdInverseAtPosition(self, pos, src=None):
        return self.cd_inverse

    def pixscale_at(self, x, y):
        return self.pixscale

    def pixel_scale(self):
        return self.pixscale

    def shifted(self, x, y):
        return self.copy()


class PixelWCS(NullWCS):
    """NullWCS whose position parameters are in pixels for derivatives too:
    d(pixel)/d(position) = I.  (An extension: the reference's NullWCS scales
    position derivatives by cd^-1; the batched configurations fit pixel
    positions with this WCS.)"""

    def cdInverseAtPosition(self, pos, src=None):
        return np.eye(2)

    def pixelDerivsToPositionDerivs(self, pos, src, counts0, patch0, patchdx, patchdy):
        derivs = []
        cdi = np.eye(2)
        for i, pname in pos.getThawedParamIndicesAndNames():
            d = (patchdx * cdi[0, i] + patchdy * cdi[1, i]) * counts0
            d.setName('d(ptsrc)/d(pos.%s)' % pname)
            derivs.append(d)
        return derivs


class PixPos(ParamList):
    """Pixel position (tractor/wcs.py:373-390): steps 0.1, maxstep 1 px."""

    @staticmethod
    def getNamedParams():
        return dict(x=0, y=1)

    def __init__(self, *args):
        super(PixPos, self).__init__(*args)
        self.stepsizes = [0.1, 0.1]
        self.maxstep = [1., 1.]

    def __str__(self):
        return 'pixel (%.2f, %.2f)' % (self.x, self.y)

    def __repr__(self):
        return 'PixPos[%s, %s]' % (repr(self.x), repr(self.y))

    def getDimension(self):
        return 2


class Flux(ScalarParam):
    """One linear flux (tractor/brightness.py:14-23)."""

    def __str__(self):
        return 'Flux: %g' % self.val

    def __repr__(self):
        return 'Flux(%r)' % self.val

    def __mul__(self, factor):
        return self.__class__(self.val * factor)

    def getFlux(self, band=None):
        return self.val


class Fluxes(ParamList):
    """Per-band fluxes (tractor/brightness.py:24-162): bands ordered by the
    `order` kwarg, else sorted by name."""

    def __init__(self, **kwargs):
        keys = kwargs.pop('order', None)
        if keys is None:
            keys = sorted(kwargs.keys())
        keys = list(keys)
        assert set(kwargs.keys()) == set(keys)
        super(Fluxes, self).__init__(*[kwargs[k] for k in keys])
        self.order = keys
        self.addNamedParams(**dict((k, i) for i, k in enumerate(keys)))
        self.stepsizes = [1.] * len(keys)

    def copy(self):
        c = self.__class__(order=self.order, **dict(zip(self.order, self.vals)))
        c.liquid = list(self.liquid)
        return c

    def __mul__(self, factor):
        return self.__class__(order=self.order, **dict((b, self.getFlux(b) * factor)
                                                       for b in self.order))

    def getParamNames(self):
        return [b for b, l in zip(self.order, self.liquid) if l]

    def getFlux(self, band):
        return self.vals[self.order.index(band)]

    def setFlux(self, band, val):
        self.vals[self.order.index(band)] = val

    def getBand(self, band):
        return self.getFlux(band)

    def setBand(self, band, val):
        self.setFlux(band, val)


class NanoMaggies(Fluxes):
    """Fluxes in nanomaggies (tractor/brightness.py:163-238)."""

    def __str__(self):
        ss = []
        for b in self.order:
            f = self.getFlux(b)
            ss.append('%s=(flux %.3g)' % (b, f) if f <= 0 else '%s=%.3f' % (b, self.getMag(b)))
        return getClassName(self) + ': ' + ', '.join(ss)

    __repr__ = __str__

    def getMag(self, band):
        return NanoMaggies.nanomaggiesToMag(self.getFlux(band))

    @staticmethod
    def magToNanomaggies(mag):
        return 10. ** ((mag - 22.5) / -2.5)

    @staticmethod
    def nanomaggiesToMag(nmgy):
        return -2.5 * (np.log10(nmgy) - 9)


class NullPhotoCal(BaseParams):
    """counts = brightness value (tractor/brightness.py:306-313)."""

    def brightnessToCounts(self, brightness):
        return brightness.getValue() if hasattr(brightness, 'getValue') else brightness.getParams()[0]

    def hashkey(self):
        return ('NullPhotoCal',)

    def copy(self):
        return NullPhotoCal()


class LinearPhotoCal(ScalarParam):
    """counts = scale * flux[band] (tractor/brightness.py:316-345)."""

    def __init__(self, scale, band=None):
        super(LinearPhotoCal, self).__init__(scale)
        self.band = band

    def copy(self):
        return LinearPhotoCal(self.val, self.band)

    def hashkey(self):
        return ('LinearPhotoCal', self.band, self.val)

    @property
    def scale(self):
        return self.val

    def getScale(self):
        return self.val

    def brightnessToCounts(self, brightness):
        if self.band is None:
            counts = brightness.getValue() if hasattr(brightness, 'getValue') else \
                brightness.getParams()[0]
        else:
            counts = brightness.getFlux(self.band)
        if counts == 0.:
            return 0.
        return counts * self.val


class NullSky(BaseParams):
    def addTo(self, img, scale=1.):
        pass

    def getValue(self):
        return 0.

    def copy(self):
        return NullSky()


class ConstantSky(ScalarParam):
    """Constant sky level (tractor/sky.py:14-56)."""

    def addTo(self, img, scale=1.):
        if self.val == 0:
            return
        img += self.val * scale

    def getParamDerivatives(self, tractor, img, srcs):
        p = Patch(0, 0, np.ones(img.shape, np.float32))
        p.setName('dsky')
        return [p]

    def getParamNames(self):
        return ['sky']
