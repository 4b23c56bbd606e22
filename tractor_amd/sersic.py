"""Sersic galaxies (tractor/sersic.py:25-500): the Sersic profile as a
mixture of Gaussians interpolated in the Sersic index, a Sersic-index
parameter, and the galaxy whose unit-flux patches go through the same device
renders as the exp and dev galaxies (FFT path, or the pixel-space MoG path
for the wide components of a hybrid PSF)."""
import functools

import numpy as np

from . import mixture_profiles as mp
from .galaxy import HoggGalaxy
from .sersic_table import CORES, RANGES
from .utils import ScalarParam


class SersicMixture(object):
    """SersicMixture (sersic.py:25-341).  Within each tabulated range of the
    index, every component's amplitude and log variance is a spline through
    the range's nodes (cubic; linear for three nodes or fewer).  Where two
    ranges overlap their mixtures are concatenated with a linear ramp of
    weights; amplitudes are normalised to 1; above n = 1 a zero-variance
    core carries the central flux the softened profile leaves out (CORES,
    cubic spline); components come sorted by variance.  Indices outside the
    table clamp to its ends."""
    _instance = None

    @staticmethod
    def getProfile(sindex):
        if SersicMixture._instance is None:
            SersicMixture._instance = SersicMixture()
        return SersicMixture._instance._profile(float(sindex))

    def __init__(self):
        from scipy.interpolate import InterpolatedUnivariateSpline as Spline
        self.core = Spline([c[0] for c in CORES], [c[1] for c in CORES], k=3)
        self.ranges = []
        for lo, hi, nodes in RANGES:
            s = [nd[0] for nd in nodes]
            k = 3 if len(nodes) > 3 else 1
            K = len(nodes[0][1])
            amp = [Spline(s, [nd[1][i] for nd in nodes], k=k) for i in range(K)]
            lvar = [Spline(s, [np.log(nd[2][i]) for nd in nodes], k=k) for i in range(K)]
            self.ranges.append((lo, hi, amp, lvar))
        self.lowest = self.ranges[0][0]
        self.highest = self.ranges[-1][1]

    @staticmethod
    def _eval(funcs, n):
        return np.array([float(f(n)) for f in funcs])

    @functools.lru_cache(maxsize=16)
    def _profile(self, n):
        if n <= self.lowest:
            found, n = [self.ranges[0]], self.lowest
        elif n >= self.highest:
            found, n = [self.ranges[-1]], self.highest
        else:
            found = [r for r in self.ranges if r[0] <= n < r[1]]
        if len(found) == 2:
            (lo0, hi0, a0, v0), (lo1, hi1, a1, v1) = found
            frac = (n - lo1) / (hi0 - lo1)
            amp0 = self._eval(a0, n)
            amp1 = self._eval(a1, n)
            amp = np.append((1. - frac) * amp0 / amp0.sum(), frac * amp1 / amp1.sum())
            var = np.exp(self._eval(v0 + v1, n))
        else:
            _, _, a, v = found[0]
            amp = self._eval(a, n)
            amp = amp / amp.sum()
            var = np.exp(self._eval(v, n))
        if n > 1.:
            core = float(self.core(n))
            amp = np.append(amp * (1. - core) / amp.sum(), core)
            var = np.append(var, 0.)
        order = np.argsort(var)
        return mp.MixtureOfGaussians(amp[order], np.zeros((len(amp), 2)), var[order])


class SersicIndex(ScalarParam):
    """SersicIndex (sersic.py:344-359): step 0.01, bounds [0.29, 6.3], at
    most 0.1 per step; zero prior inside the bounds, -inf outside."""
    stepsize = 0.01

    def __init__(self, val=0):
        super(SersicIndex, self).__init__(val=val)
        self.lower = 0.29
        self.upper = 6.3
        self.maxstep = 0.1

    def isLegal(self):
        return self.lower <= self.val <= self.upper

    def getLogPrior(self):
        return 0. if self.isLegal() else -np.inf

    def copy(self):
        return SersicIndex(self.val)


class SersicGalaxy(HoggGalaxy):
    """SersicGalaxy (sersic.py:362-500): position, brightness, shape and
    Sersic index; half-size 8 r_e like the dev profile; the index derivative
    by a forward difference of the unit-flux render (step 0.01, backwards
    when the step would pass the upper bound)."""
    nre = 8.

    @staticmethod
    def getNamedParams():
        return dict(pos=0, brightness=1, shape=2, sersicindex=3)

    def __init__(self, pos, brightness, shape, sersicindex, **kwargs):
        self.nre = SersicGalaxy.nre
        super(SersicGalaxy, self).__init__(pos, brightness, shape, sersicindex)

    def getName(self):
        return 'SersicGalaxy'

    def __str__(self):
        return super(SersicGalaxy, self).__str__() + ', Sersic index %.3f' % self.sersicindex.val

    def getProfile(self):
        return SersicMixture.getProfile(self.sersicindex.val)

    def copy(self):
        c = SersicGalaxy(self.pos.copy(), self.brightness.copy(), self.shape.copy(),
                         self.sersicindex.copy())
        c.liquid = list(self.liquid)
        return c

    def _extra_columns(self, plan, img, px, py, modelMask, counts, kwargs):
        """The Sersic-index column (sersic.py:402-442): a forward difference
        of the unit render on the caller's modelMask, the step reversed when
        it would pass the upper bound."""
        if self.isParamFrozen('sersicindex'):
            return []
        si = self.sersicindex
        names, vals, ups = si.getParamNames(), si.getParams(), si.getUpperBounds()
        cols = []
        for i, step in enumerate(si.getStepSizes()):
            if vals[i] + step > ups[i]:
                step = -step
            r = self._stepped(plan, img, si, i, step, px=px, py=py, modelMask=modelMask, **kwargs)
            cols.append(('diff', r, counts / step, 'd(%s)/d(%s)' % (self.dname, names[i])))
        return cols


def _pieces(nodes, vals, k):
    """Breakpoints and per-piece cubic coefficients (highest power first, in
    powers of n - breakpoint) of the interpolating spline of degree k through
    (nodes, vals), with the zero-length pieces of repeated end knots dropped."""
    from scipy.interpolate import PPoly, splrep
    pp = PPoly.from_spline(splrep(nodes, vals, k=k, s=0))
    keep = np.nonzero(np.diff(pp.x) > 0)[0]
    br = np.append(pp.x[keep], pp.x[keep[-1] + 1])
    c = np.zeros((len(keep), 4))
    c[:, 3 - k:] = pp.c[:, keep].T
    return br, c


@functools.lru_cache(maxsize=1)
def fitter_table():
    """The Sersic mixture (SersicMixture above) as the flat table of piecewise
    cubics that trh_fitter_set_sersic takes: R, then per range lo, hi, K, m,
    the m + 1 breakpoints and, per component, m amplitude and m log-variance
    pieces of 4 coefficients; then the core fraction's m, breakpoints and
    pieces.  The splines are the same interpolants as SersicMixture's."""
    out = [float(len(RANGES))]
    for lo, hi, nodes in RANGES:
        s = [nd[0] for nd in nodes]
        k = 3 if len(nodes) > 3 else 1
        K = len(nodes[0][1])
        comps = []
        br0 = None
        for i in range(K):
            for vals in ([nd[1][i] for nd in nodes], [np.log(nd[2][i]) for nd in nodes]):
                br, c = _pieces(s, vals, k)
                if br0 is None:
                    br0 = br
                assert np.array_equal(br, br0)
                comps.append(c.ravel())
        out += [lo, hi, float(K), float(len(br0) - 1)] + list(br0)
        for c in comps:
            out += list(c)
    br, c = _pieces([c[0] for c in CORES], [c[1] for c in CORES], 3)
    out += [float(len(br) - 1)] + list(br) + list(c.ravel())
    return np.array(out, np.float64)
