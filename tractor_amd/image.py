"""Image: pixels, inverse error and calibration Params (tractor/image.py:6-92)."""
import numpy as np

from .utils import MultiParams


class Image(MultiParams):
    """Image(data, invvar=None, inverr=None, psf, wcs, sky, photocal, name)."""

    def __init__(self, data=None, invvar=None, inverr=None, psf=None, wcs=None, sky=None,
                 photocal=None, name=None, time=None, **kwargs):
        from .calib import NullWCS, NullSky, NullPhotoCal
        self.data = data
        if inverr is not None:
            self.inverr = inverr
        elif invvar is not None:
            self.setInvvar(invvar)
        self.name = name
        self.zr = kwargs.pop('zr', None)
        self.time = time
        if wcs is None:
            wcs = NullWCS()
        if sky is None:
            sky = NullSky()
        if photocal is None:
            photocal = NullPhotoCal()
        self.modelMinval = 0.
        super(Image, self).__init__(psf, wcs, photocal, sky)

    def __str__(self):
        return 'Image ' + str(self.name)

    @staticmethod
    def getNamedParams():
        return dict(psf=0, wcs=1, photocal=2, sky=3)

    def copy(self):
        return self.__class__(data=self.data.copy(), inverr=self.inverr.copy(),
                              wcs=self.wcs.copy(), psf=self.psf.copy(), sky=self.sky.copy(),
                              photocal=self.photocal.copy(), name=self.name, time=self.time)

    def subimage(self, x0, x1, y0, y1):
        slc = (slice(y0, y1), slice(x0, x1))
        sub = Image(data=self.data[slc].copy(), inverr=self.inverr[slc].copy(),
                    wcs=self.wcs.shifted(x0, y0), psf=self.psf.getShifted(x0, y0),
                    sky=self.sky, photocal=self.photocal.copy())
        sub.name = self.name
        return sub

    def getParamDerivatives(self, tractor, srcs):
        derivs = []
        for s in self._getActiveSubs():
            if hasattr(s, 'getParamDerivatives'):
                sd = s.getParamDerivatives(tractor, self, srcs)
                assert len(sd) == s.numberOfParams()
                derivs.extend(sd)
            else:
                derivs.extend([False] * s.numberOfParams())
        return derivs

    @property
    def shape(self):
        return self.getShape()

    def getShape(self):
        return self.data.shape

    def getModelShape(self):
        return self.getShape()

    def numberOfPixels(self):
        H, W = self.shape
        return H * W

    def getInvError(self):
        return self.inverr

    def getInvvar(self):
        return self.inverr ** 2

    def setInvvar(self, iv):
        self.inverr = np.sqrt(iv)

    def getImage(self):
        return self.data

    def setImage(self, img):
        self.data = img

    def getPsf(self):
        return self.psf

    def getWcs(self):
        return self.wcs

    def getSky(self):
        return self.sky

    def getPhotoCal(self):
        return self.photocal

    def setPsf(self, psf):
        self.psf = psf

    def setWcs(self, wcs):
        self.wcs = wcs

    def setSky(self, sky):
        self.sky = sky

    def setPhotoCal(self, pc):
        self.photocal = pc
