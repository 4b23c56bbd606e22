"""Galaxy shape parameterisations (tractor/ellipses.py:15-264)."""
import math

import numpy as np

from .utils import ParamList
from .mixture_profiles import e_basis


class EllipseE(ParamList):
    """(re [arcsec], e1, e2): steps 0.01, bounds re>=0, |e_i|<=1,
    maxstep [1, 0.1, 0.1] (ellipses.py:15-163)."""

    @staticmethod
    def getName():
        return 'EllipseE'

    @staticmethod
    def getNamedParams():
        return dict(re=0, e1=1, e2=2)

    def __init__(self, *args, **kwargs):
        super(EllipseE, self).__init__(*args, **kwargs)
        self.stepsizes = [0.01] * 3
        self.lowers = [0., -1., -1.]
        self.uppers = [None, 1., 1.]
        self.maxstep = [1., 0.1, 0.1]

    @staticmethod
    def fromRAbPhi(r, ba, phi):
        ab = 1. / ba
        e = (ab - 1) / (ab + 1)
        angle = math.radians(2. * (-phi))
        return EllipseE(r, e * math.cos(angle), e * math.sin(angle))

    @property
    def e(self):
        return math.hypot(self.e1, self.e2)

    @property
    def theta(self):
        return math.atan2(self.e2, self.e1) / 2.

    @property
    def ab(self):
        e = self.e
        return (1. - e) / (1. + e)

    def __repr__(self):
        return 're=%g, e1=%g, e2=%g' % (self.re, self.e1, self.e2)

    def __str__(self):
        return self.getName() + ': ' + repr(self)

    def isLegal(self):
        return ((self.e1 ** 2 + self.e2 ** 2) < 1.) and (self.re >= 0.)

    def getRaDecBasis(self):
        """r_e unit vectors -> degrees (ellipses.py:126-155)."""
        return e_basis(self.re, self.e, self.theta)

    def getCovariance(self):
        G = self.getRaDecBasis() * 3600.
        return G @ G.T


class EllipseESoft(EllipseE):
    """(log re, ee1, ee2), e = 1 - exp(-|ee|); no bounds (ellipses.py:165-264)."""

    @staticmethod
    def getName():
        return 'EllipseESoft'

    @staticmethod
    def getNamedParams():
        return dict(logre=0, ee1=1, ee2=2)

    def __init__(self, *args, **kwargs):
        super(EllipseESoft, self).__init__(*args, **kwargs)
        self.lowers = [None, None, None]
        self.uppers = [None, None, None]

    @staticmethod
    def fromEllipseE(ell, maxe=0.999999):
        e = min(ell.e, maxe)
        esoft = -math.log(1. - e)
        if e == 0:
            return EllipseESoft(math.log(ell.re), 0., 0.)
        return EllipseESoft(math.log(ell.re), ell.e1 / e * esoft, ell.e2 / e * esoft)

    def __repr__(self):
        return 'log r_e=%.3f, ee1=%.3f, ee2=%.3f' % (self.logre, self.ee1, self.ee2)

    @property
    def re(self):
        return math.exp(self.logre)

    @property
    def e(self):
        return 1. - math.exp(-math.hypot(self.ee1, self.ee2))

    @property
    def theta(self):
        return math.atan2(self.ee2, self.ee1) / 2.

    def isLegal(self):
        return True
