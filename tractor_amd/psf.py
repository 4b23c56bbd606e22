"""PSF models (tractor/psf.py): PixelizedPSF (Lanczos-shifted stamp; FFT
spectrum for galaxies), GaussianMixturePSF, NCircularGaussianPSF,
HybridPixelizedPSF.  Pixels are rendered by libtractor_hip."""
import numpy as np

from .utils import BaseParams, ParamList, MultiParams
from .mixture_profiles import MixtureOfGaussians, mixture_to_patch, plan_mixture
from .patch import Patch
from .plan import Pending, RenderPlan, ready, resolve
from . import render as R


def lanczos_shift_image(img, dx, dy, inplace=False, force_python=False):
    """tractor/psf.py:26-58 -- Lanczos-3 shift on the device."""
    plan = RenderPlan()
    s = plan.shift(img, dx, dy)
    out = plan.run().result(s)
    if inplace:
        img[:, :] = out
    return out


def central_differences(p):
    """(d/dx, d/dy) of a patch by central differences, negated (the
    derivative with respect to the source position), with a one-pixel zero
    border: the rule of PixelizedPSF.getPointSourcePatch (tractor/psf.py:
    196-207) and Galaxy.getParamDerivatives (tractor/galaxy.py:168-172)."""
    gx = np.zeros_like(p)
    gy = np.zeros_like(p)
    gx[1:-1, 1:-1] = 0.5 * (p[1:-1, :-2] - p[1:-1, 2:])
    gy[1:-1, 1:-1] = 0.5 * (p[:-2, 1:-1] - p[2:, 1:-1])
    return gx, gy


def plan_point_patch(psf, plan, px, py, **kwargs):
    """psf's point-source patch as a Pending in `plan`; PSF classes without
    a planner render at once (getPointSourcePatch)."""
    fn = getattr(psf, '_plan_point', None)
    if fn is None:
        r = psf.getPointSourcePatch(px, py, **kwargs)
        if r is None:
            return ready(None)
        p0 = r[0] if isinstance(r, tuple) else r
        return ready(r, (p0.x0, p0.y0, p0.patch.shape[1], p0.patch.shape[0]),
                     has_derivs=isinstance(r, tuple))
    return fn(plan, px, py, **kwargs)


class HybridPSF(object):
    """Marker: galaxies render wide components with the MoG (psf.py:65-66)."""
    pass


class PixelizedPSF(BaseParams):
    """PSF from an odd-sized image stamp (tractor/psf.py:68-367), sampling 1."""

    def __init__(self, img, sampling=1., Lorder=3):
        img = np.asarray(img).astype(np.float32)
        H, W = img.shape
        assert (H % 2) == 1 and (W % 2) == 1
        if sampling != 1.:
            raise NotImplementedError('PixelizedPSF sampling != 1 is not supported')
        self.img = img
        self.radius = np.hypot(H / 2., W / 2.)
        self.H, self.W = H, W
        self.Lorder = Lorder
        self.sampling = sampling
        self.fftcache = {}

    def __str__(self):
        return 'PixelizedPSF'

    def __getstate__(self):
        d = self.__dict__.copy()
        d.pop('_trh_spectra', None)
        d['fftcache'] = {}
        return d

    def clear_cache(self):
        self.fftcache = {}
        self.__dict__.pop('_trh_spectra', None)

    @property
    def shape(self):
        return (self.H, self.W)

    def hashkey(self):
        return ('PixelizedPSF', tuple(self.img.ravel()))

    def copy(self):
        return self.__class__(self.img.copy())

    def getShifted(self, x0, y0):
        return self

    def constantPsfAt(self, x, y):
        return self

    def getRadius(self):
        return self.radius

    def getImage(self, px, py):
        return self.img

    def getPointSourcePatch(self, px, py, minval=0., modelMask=None, radius=None, derivs=False,
                            **kwargs):
        """tractor/psf.py:137-207: the Lanczos-shifted stamp, on the
        modelMask if given (None without overlap), with (dx, dy) patches by
        central differences when derivs and a modelMask are given."""
        return resolve(self._plan_point, px, py, minval=minval, modelMask=modelMask,
                       radius=radius, derivs=derivs)

    def _cropped(self, radius):
        # the stamp trimmed to `radius` about its centre (psf.py:149-156),
        # cached per radius so its device copy is made once
        R_ = int(np.ceil(radius))
        crops = self.__dict__.setdefault('_crops', {})
        if R_ not in crops:
            H, W = self.img.shape
            cx, cy = W // 2, H // 2
            sub = self.img[max(cy - R_, 0):min(cy + R_ + 1, H - 1),
                           max(cx - R_, 0):min(cx + R_ + 1, W - 1)]
            c = PixelizedPSF.__new__(PixelizedPSF)
            c.img = np.ascontiguousarray(sub, np.float32)
            c.H, c.W = c.img.shape
            crops[R_] = c
        return crops[R_]

    def _plan_point(self, plan, px, py, minval=0., modelMask=None, radius=None, derivs=False,
                    **kwargs):
        src = self if radius is None else self._cropped(radius)
        stamp = src.img
        H, W = stamp.shape
        ix = int(round(float(px)))
        iy = int(round(float(py)))
        x0, y0 = ix - W // 2, iy - H // 2
        if modelMask is None:
            slot = plan.shift(stamp, px - ix, py - iy)
            return Pending((x0, y0, W, H), lambda: Patch(x0, y0, plan.result(slot)))
        mm = (modelMask.x0, modelMask.y0, modelMask.w, modelMask.h)
        J = R.point_job(px, py, stamp.shape, mm)
        if J is None:
            return ready(None, has_derivs=derivs)
        slot = plan.point(src, J, modelMask.w, modelMask.h)

        def done():
            p = plan.result(slot)
            patch = Patch(mm[0], mm[1], p)
            if not derivs:
                return patch
            gx, gy = central_differences(p)
            return patch, Patch(mm[0], mm[1], gx), Patch(mm[0], mm[1], gy)
        return Pending(mm, done, has_derivs=derivs)

    def getFourierTransformSize(self, radius):
        return R.fft_size(radius)

    def getFourierTransform(self, px, py, radius):
        """(P, (cx, cy), (pH, pW), (v, w)) of psf.py:249-280, with P the
        device half spectrum (torch complex64 view) instead of NumPy."""
        sz = self.getFourierTransformSize(radius)
        if sz not in self.fftcache:
            from .plan import _spectra_of
            spec, cx, cy = _spectra_of(self).get(sz)
            self.fftcache[sz] = (spec, (cx, cy), (sz, sz),
                                 (np.fft.rfftfreq(sz), np.fft.fftfreq(sz)))
        return self.fftcache[sz]


class GaussianMixturePSF(BaseParams):
    """PSF as a mixture of Gaussians (tractor/psf.py:369-528):
    GaussianMixturePSF(amp, mean, var) or (a0.., mx0, my0, .., vxx, vyy, vxy ...)."""

    def __init__(self, *args):
        if len(args) == 3 and not np.isscalar(args[0]):
            amp, mean, var = args
            self.mog = MixtureOfGaussians(np.atleast_1d(amp), np.atleast_2d(mean), var)
        else:
            n = len(args)
            K = n // 6
            a = np.array(args, float)
            amp = a[:K]
            mean = a[K:3 * K].reshape(K, 2)
            v = a[3 * K:].reshape(K, 3)
            var = np.zeros((K, 2, 2))
            var[:, 0, 0] = v[:, 0]
            var[:, 1, 1] = v[:, 1]
            var[:, 0, 1] = var[:, 1, 0] = v[:, 2]
            self.mog = MixtureOfGaussians(amp, mean, var)
        self.radius = 25
        K = self.mog.K
        self.stepsizes = [0.01] * K + [0.01] * (K * 2) + [0.1] * (K * 3)

    def __str__(self):
        return ('GaussianMixturePSF: amps=' + str(tuple(self.mog.amp.ravel())) +
                ', means=' + str(tuple(self.mog.mean.ravel())) +
                ', var=' + str(tuple(self.mog.var.ravel())))

    def copy(self):
        return GaussianMixturePSF(self.mog.amp.copy(), self.mog.mean.copy(), self.mog.var.copy())

    @staticmethod
    def fromStamp(stamp, N=3, P0=None, xy0=None, alpha=0., emsteps=1000, v2=False,
                  approx=1e-30, v3=False):
        """EM fit of an N-component mixture to a PSF stamp on the device
        (GaussianMixturePSF.fromStamp, tractor/psf.py:537-597 with
        em_fit_2d_reg, emfit.i:274; trh_em_fit_psf).  P0 = (w, mu, var) or
        fitpsf.em_init_params's start; xy0 = the stamp's origin (default
        minus half its size).  The stamp is normalised to sum 1 and clipped
        at 0 first."""
        if v2 or v3:
            raise NotImplementedError('fromStamp: only the em_fit_2d_reg fit is built')
        return GaussianMixturePSF(*fit_mog_to_stamps([stamp], N, P0, xy0, alpha, emsteps)[0])

    def hashkey(self):
        return ('GaussianMixturePSF', tuple(self.mog.amp), tuple(self.mog.mean.ravel()),
                tuple(self.mog.var.ravel()))

    def getShifted(self, x0, y0):
        return self

    def constantPsfAt(self, x, y):
        return self

    def getMixtureOfGaussians(self, px=None, py=None):
        return self.mog

    def getRadius(self):
        return self.radius

    def getPointSourcePatch(self, px, py, minval=0., radius=None, derivs=False, minradius=None,
                            modelMask=None, clipExtent=None, **kwargs):
        """tractor/psf.py:458-528: exact grid, masked (with derivatives) or
        ring-growth (minval > 0) evaluation."""
        return resolve(self._plan_point, px, py, minval=minval, radius=radius, derivs=derivs,
                       minradius=minradius, modelMask=modelMask, clipExtent=clipExtent)

    def _box(self, px, py, r, minval):
        # the evaluation box about (px, py) of radius r; r None: where every
        # component falls below minval (an overestimate)
        if r is None:
            r = 0.
            for v in self.mog.var:
                r2 = (v[0, 0] + v[1, 1]) * -2. * np.log(minval * 2. * np.pi * np.linalg.det(v))
                r = max(r, np.sqrt(r2)) if r2 > 0 else r
            r = int(np.ceil(r))
        return (int(np.floor(px - r)), int(np.ceil(px + r)) + 1,
                int(np.floor(py - r)), int(np.ceil(py + r)) + 1)

    def _plan_point(self, plan, px, py, minval=0., radius=None, derivs=False, minradius=None,
                    modelMask=None, clipExtent=None, **kwargs):
        g = self.mog
        if modelMask is not None:
            mm = modelMask
            ext = (mm.x0, mm.y0, mm.w, mm.h)
            if mm.mask is None:
                slot = plan.mog(g.amp, g.mean, g.var, mm.x0, mm.y0, mm.w, mm.h, px, py)
                return Pending(ext, lambda: Patch(mm.x0, mm.y0, plan.result(slot)))
            slot = plan.mog(g.amp, g.mean, g.var, mm.x0, mm.y0, mm.w, mm.h, px, py, mask=mm.mask,
                            derivs=derivs, masked_rules=True)
            if derivs:
                return Pending(ext, lambda: tuple(Patch(mm.x0, mm.y0, q)
                                                  for q in plan.result_derivs(slot)),
                               has_derivs=True)
            return Pending(ext, lambda: Patch(mm.x0, mm.y0, plan.result(slot)))
        minval = 0. if minval is None else minval
        approx = minval > 0. or minradius is not None
        if approx:
            r = radius if radius is not None else self.radius
        else:
            r = self.getRadius() if radius is None else radius
        x0, x1, y0, y1 = self._box(px, py, r, minval)
        if clipExtent is not None:
            xl, xh, yl, yh = clipExtent
            x0, x1, y0, y1 = max(x0, xl), min(x1, xh), max(y0, yl), min(y1, yh)
        if x0 >= x1 or y0 >= y1:
            return ready(None)
        if approx:
            # ring growth (minval > 0) is evaluated at once; its window is the
            # evaluated region, known only afterwards
            kwa = {} if minradius is None else dict(minradius=minradius)
            r = g.evaluate_grid_approx3(x0, x1, y0, y1, px, py, minval, derivs=derivs, **kwa)
            if r is None:
                return ready(None)
            p0 = r[0] if derivs else r
            return ready(r, (p0.x0, p0.y0, p0.patch.shape[1], p0.patch.shape[0]),
                         has_derivs=derivs)
        slot = plan.mog(g.amp, g.mean, g.var, x0, y0, x1 - x0, y1 - y0, px, py)
        return Pending((x0, y0, x1 - x0, y1 - y0), lambda: Patch(x0, y0, plan.result(slot)))


class NCircularGaussianPSF(MultiParams):
    """N concentric circular Gaussians (tractor/psf.py:792-929)."""

    def __init__(self, sigmas, weights):
        assert len(sigmas) == len(weights)
        ps = ParamList(*sigmas)
        ps.stepsizes = [0.01] * len(sigmas)
        pw = ParamList(*weights)
        pw.stepsizes = [0.01] * len(weights)
        super(NCircularGaussianPSF, self).__init__(ps, pw)
        self.minradius = 1.

    @staticmethod
    def getNamedParams():
        return dict(sigmas=0, weights=1)

    def __str__(self):
        return ('NCircularGaussianPSF: sigmas [ ' + ', '.join('%.3f' % s for s in self.mysigmas) +
                ' ], weights [ ' + ', '.join('%.3f' % w for w in self.myweights) + ' ]')

    def copy(self):
        return NCircularGaussianPSF(list(self.mysigmas), list(self.myweights))

    def getShifted(self, x0, y0):
        return self

    def constantPsfAt(self, x, y):
        return self

    @property
    def mysigmas(self):
        return self.sigmas.vals

    @property
    def myweights(self):
        return self.weights.vals

    def getMixtureOfGaussians(self, px=None, py=None):
        """getCircularMog (psf.py:931-938): float32 amps and variances."""
        K = len(self.mysigmas)
        amps = np.array(self.myweights).astype(np.float32)
        var = np.zeros((K, 2, 2), np.float32)
        for k, s in enumerate(self.mysigmas):
            var[k, 0, 0] = var[k, 1, 1] = s ** 2
        return MixtureOfGaussians(amps, np.zeros((K, 2), np.float32), var, quick=True)

    def getNSigma(self):
        return 5.

    def getRadius(self):
        if hasattr(self, 'radius'):
            return self.radius
        return max(self.minradius, max(self.mysigmas) * self.getNSigma())

    def getPointSourcePatch(self, px, py, minval=0., radius=None, modelMask=None,
                            clipExtent=None, **kwargs):
        """tractor/psf.py:898-929 via mixture_to_patch (exact grid or ring growth)."""
        return resolve(self._plan_point, px, py, minval=minval, radius=radius,
                       modelMask=modelMask, clipExtent=clipExtent)

    def _plan_point(self, plan, px, py, minval=0., radius=None, modelMask=None, clipExtent=None,
                    **kwargs):
        if modelMask is not None:
            x0, x1, y0, y1 = modelMask.extent
        else:
            ix, iy = int(round(px)), int(round(py))
            rad = int(np.ceil(self.getRadius())) if radius is None else radius
            x0, x1, y0, y1 = ix - rad, ix + rad + 1, iy - rad, iy + rad + 1
        if clipExtent is not None:
            xl, xh, yl, yh = clipExtent
            x0, x1, y0, y1 = max(x0, xl), min(x1, xh), max(y0, yl), min(y1, yh)
        mix = self.getMixtureOfGaussians()
        # the reference adds (px, py) into the float32 means (psf.py:921-925)
        mean = mix.mean.copy()
        mean[:, 0] += px
        mean[:, 1] += py
        mix = MixtureOfGaussians(mix.amp, mean, mix.var, quick=True)
        return plan_mixture(plan, mix, x0, x1, y0, y1, minval=minval,
                            exactExtent=(modelMask is not None))


class HybridPixelizedPSF(HybridPSF):
    """PixelizedPSF + a Gaussian-mixture approximation (tractor/psf.py:600-655).
    gauss=None fits an N-component mixture to the stamp at (cx, cy) on the
    device (GaussianMixturePSF.fromStamp)."""

    def __init__(self, pix, gauss=None, N=2, cx=0., cy=0.):
        if gauss is None:
            gauss = GaussianMixturePSF.fromStamp(pix.getImage(cx, cy), N=N)
        self.__dict__['pix'] = pix
        self.__dict__['gauss'] = gauss

    def __str__(self):
        return 'HybridPixelizedPSF: Pix %s' % str(self.pix)

    def copy(self):
        return HybridPixelizedPSF(self.pix.copy(), self.gauss.copy())

    def getMixtureOfGaussians(self, **kwargs):
        return self.gauss.getMixtureOfGaussians(**kwargs)

    def getShifted(self, dx, dy):
        return HybridPixelizedPSF(self.pix.getShifted(dx, dy), self.gauss)

    def constantPsfAt(self, x, y):
        return HybridPixelizedPSF(self.pix.constantPsfAt(x, y), self.gauss)

    def __getattr__(self, name):
        if name in ('pix', 'gauss') or name.startswith('__'):
            raise AttributeError(name)
        return getattr(self.__dict__['pix'], name)

    def __getstate__(self):
        return (self.gauss, self.pix)

    def __setstate__(self, state):
        self.__dict__['gauss'], self.__dict__['pix'] = state


def fit_mog_to_stamps(stamps, N=3, P0=None, xy0=None, alpha=0., emsteps=1000):
    """Batched EM fit (trh_em_fit_psf) of an N-component mixture of Gaussians
    to each stamp (all of one shape): the stamp normalised to sum 1 and
    clipped at 0, start P0 = (w, mu, var) or equal weights, zero means and
    variances I (k + 1) (fitpsf.em_init_params, fitpsf.py:72-86).  Returns
    [(amp, mean, var)] per stamp."""
    import torch
    from . import _lib
    from ._lib import check, ptr
    st = np.stack([np.asarray(s, np.float64) for s in stamps])
    n, NY, NX = st.shape
    st = st / st.sum(axis=(1, 2), keepdims=True)
    st = np.maximum(st, 0.)
    if P0 is not None:
        w, mu, var = (np.asarray(a, np.float64) for a in P0)
    else:
        w = np.ones(N) / N
        mu = np.zeros((N, 2))
        var = np.array([np.eye(2) * (k + 1) for k in range(N)])
    K = len(w)
    x0, y0 = (-(NX // 2), -(NY // 2)) if xy0 is None else xy0
    dev = torch.device('cuda', 0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)
    d_st = t(st)
    d_amp = t(np.tile(w, (n, 1)))
    d_mean = t(np.tile(mu.reshape(1, K, 2), (n, 1, 1)))
    d_var = t(np.tile(var.reshape(1, K, 2, 2), (n, 1, 1, 1)))
    status = torch.zeros(n, dtype=torch.int32, device=dev)
    lib = _lib.load()
    h = _lib.handle(0)
    check(lib.trh_em_fit_psf(h.value, n, ptr(d_st), NY, NX, int(x0), int(y0), K, float(alpha),
                             int(emsteps), ptr(d_amp), ptr(d_mean), ptr(d_var), ptr(status),
                             None), 'em_fit_psf')
    if np.any(status.cpu().numpy() != 0):
        raise ValueError('EM fit: a variance lost positive definiteness')
    a, m, v = d_amp.cpu().numpy(), d_mean.cpu().numpy(), d_var.cpu().numpy()
    return [(a[i], m[i], v[i]) for i in range(n)]
