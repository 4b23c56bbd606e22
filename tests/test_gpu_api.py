"""The drop-in API on the device, mirroring the reference's own tests:
test/test_galaxy.py (patch extents, sums, ModelMask variants, PixelizedPSF /
HybridPixelizedPSF), test/test_tractor.py (model images, lnp = -143.681,
forced photometry) and test/test_forced_phot.py (single-pixel closed forms,
Gaussian flux prior).  Every pixel comes from libtractor_hip."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import tractor_amd as T
from tractor_amd.patch import Patch, ModelMask
from tractor_amd.utils import _GaussianPriors, GaussianPriorsMixin, MultiParams


@pytest.fixture(scope='module', autouse=True)
def _need_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')


def _tim(psf, H=100, W=100):
    return T.Image(data=np.zeros((H, W), np.float32), inverr=np.ones((H, W), np.float32),
                   psf=psf, photocal=T.LinearPhotoCal(1., band='r'))


def _add(p, H=100, W=100):
    m = np.zeros((H, W), np.float32)
    p.addTo(m)
    return m


def test_galaxy_sequence():
    """test/test_galaxy.py:21-475 step by step."""
    pos0 = T.PixPos(49.5, 50.)
    bright = T.NanoMaggies(g=1., r=2.)
    shape = T.GalaxyShape(2., 0.5, 45.)
    psf0 = T.GaussianMixturePSF(1., 0., 0., 6., 6., -1.)
    tim = _tim(psf0)
    gal1 = T.ExpGalaxy(pos0, bright, shape)

    derivs = gal1.getParamDerivatives(tim)
    assert len(derivs) == 7
    for i, d in enumerate(derivs):
        assert (d is None) == (i == 2)          # r-band image: g flux derivative is None
    gal1.brightness.r = 0.
    derivs = gal1.getParamDerivatives(tim)
    for i, d in enumerate(derivs):
        assert (d is not None) == (i == 3)      # only the r flux derivative survives
    gal1.brightness.r = 100.

    p1 = gal1.getModelPatch(tim)
    assert (p1.x0, p1.y0, p1.shape) == (24, 25, (52, 53))
    assert abs(p1.patch.sum() - 100.) < 1e-3
    mod = _add(p1)

    mm = ModelMask(25, 25, 50, 50)
    p2 = gal1.getModelPatch(tim, modelMask=mm)
    assert (p2.x0, p2.y0, p2.shape) == (25, 25, (50, 50))
    mod2 = _add(p2)
    assert abs(mod2.sum() - 100.) < 1e-3
    assert np.abs(mod - mod2).max() < 1e-6

    mm3 = ModelMask(30, 29, np.ones((40, 40), bool))
    p3 = gal1.getModelPatch(tim, modelMask=mm3)
    assert (p3.x0, p3.y0, p3.shape) == (30, 29, (40, 40))
    mod3 = _add(p3)
    assert abs(mod3.sum() - 100.) < 1e-3
    assert np.abs(mod3 - mod).max() < 1e-6

    psfpatch = tim.psf.getPointSourcePatch(24., 24., modelMask=ModelMask(0, 0, 50, 50))
    pixpsf = T.PixelizedPSF(psfpatch.patch[:49, :49])
    tim.psf = pixpsf
    p4 = gal1.getModelPatch(tim)
    assert (p4.x0, p4.y0, p4.shape) == (14, 15, (72, 73))
    mod4 = _add(p4)
    assert abs(mod4.sum() - 100.) < 1e-3
    assert np.abs(mod4 - mod).max() < 2e-3

    p5 = gal1.getModelPatch(tim, modelMask=mm)
    assert (p5.x0, p5.y0, p5.shape) == (25, 25, (50, 50))
    mod5 = _add(p5)
    assert abs(mod5.sum() - 100.) < 1e-3
    assert np.abs(mod5 - mod).max() < 2e-3

    gal1.pos = T.PixPos(200, -50.)
    assert gal1.getModelPatch(tim, modelMask=mm) is None

    tim.psf = T.HybridPixelizedPSF(pixpsf, gauss=psf0)
    gal1.pos = T.PixPos(20., 25.)
    p7 = gal1.getModelPatch(tim, modelMask=mm)
    assert (p7.x0, p7.y0, p7.shape) == (25, 25, (50, 50))
    mod7 = _add(p7)
    assert abs(mod7.sum() - 1.963) < 1e-3

    # test_galaxy.py:361-379: wrap again with gauss=None, which EM-fits a
    # 2-component mixture to the stamp; the rest of the sequence uses it.
    tim.psf = T.HybridPixelizedPSF(tim.psf)
    g = tim.psf.gauss.mog
    cov = sum(g.amp[k] * (g.var[k] + np.outer(g.mean[k], g.mean[k])) for k in range(2))
    assert abs(g.amp.sum() - 1.) < 1e-6
    np.testing.assert_allclose(cov, [[6., -1.], [-1., 6.]], atol=0.1)
    p8 = gal1.getModelPatch(tim, modelMask=mm)
    assert (p8.x0, p8.y0, p8.shape) == (25, 25, (50, 50))
    assert abs(_add(p8).sum() - 1.963) < 5e-2

    gal1.pos = T.PixPos(5., 5.)
    mod9 = _add(gal1.getModelPatch(tim))
    assert abs(mod9.sum() - 94.33) < 1e-2
    assert np.sum(np.abs(mod9[50:, :])) == 0.
    assert np.sum(np.abs(mod9[:, 50:])) == 0.

    gal1.pos = pos0
    p10 = gal1.getModelPatch(tim, modelMask=ModelMask(45, 45, 10, 10))
    assert (p10.x0, p10.y0, p10.shape) == (45, 45, (10, 10))
    mod10 = _add(p10)
    mod11 = _add(gal1.getModelPatch(tim, modelMask=ModelMask(30, 30, 40, 40)))
    assert abs(mod11.sum() - 100.) < 1e-3
    assert np.abs((mod11 - mod10)[45:55, 45:55]).max() < 1e-6

    gal2 = T.DevGalaxy(pos0, bright, T.GalaxyShape(3., 0.4, 60.))
    mod12 = _add(gal2.getModelPatch(tim, modelMask=mm))
    assert abs(mod12.sum() - 99.95) < 1e-2


def test_galaxy_patch_matches_oracle():
    """Whole-patch comparison of the FFT path with the FP64 oracle."""
    from oracle import tractor_oracle as O
    psf0 = T.GaussianMixturePSF(1., 0., 0., 6., 6., -1.)
    psfpatch = psf0.getPointSourcePatch(24., 24., modelMask=ModelMask(0, 0, 50, 50))
    pix = psfpatch.patch[:49, :49].astype(np.float32)
    tim = _tim(T.PixelizedPSF(pix))
    gal = T.ExpGalaxy(T.PixPos(49.3, 50.2), T.NanoMaggies(r=100.), T.GalaxyShape(2., 0.5, 45.))
    mm = ModelMask(25, 25, 50, 50)
    p = gal.getModelPatch(tim, modelMask=mm)
    amp, var = O.exp_mixture()
    var = O.shear_mixture(var, np.eye(2) * 3600., O.galaxyshape_basis(2., 0.5, 45.))
    x0, y0, ref = O.render_galaxy_fft(100. * amp, var, 49.3, 50.2, pix, mm=(25, 25, 50, 50))
    assert (p.x0, p.y0) == (x0, y0)
    print('  galaxy patch max err / peak %.3e' % (np.abs(p.patch - ref).max() / np.abs(ref).max()),
          flush=True)
    np.testing.assert_allclose(p.patch, ref, rtol=0, atol=1e-5 * np.abs(ref).max())


def test_model_images_and_forced_phot():
    """test/test_tractor.py:53-124."""
    W, H = 100, 100
    tim1 = T.Image(data=np.zeros((H, W)), invvar=np.ones((H, W)),
                   psf=T.NCircularGaussianPSF([2.], [1.]), photocal=T.LinearPhotoCal(1.))
    tim2 = T.Image(data=np.zeros((H, W)), invvar=np.ones((H, W)),
                   psf=T.NCircularGaussianPSF([3.], [1.]), photocal=T.LinearPhotoCal(1.))
    trueflux = 100.
    star = T.PointSource(T.PixPos(W / 2, H / 2), T.Flux(trueflux))
    tr = T.Tractor([tim1, tim2], [star])
    mods = list(tr.getModelImages())
    assert abs(np.sum(mods[0]) - trueflux) < 1e-3
    assert abs(np.sum(mods[1]) - trueflux) < 1e-3
    assert abs(np.max(mods[0]) - trueflux / (2. * np.pi * 2. ** 2)) < 1e-3
    assert abs(np.max(mods[1]) - trueflux / (2. * np.pi * 3. ** 2)) < 1e-3
    lnp = tr.getLogProb()
    assert abs(lnp - -143.681) < 1e-3

    tr.freezeParam('images')
    tr.optimize()
    star.brightness.setParams([100.])
    star.freezeAllBut('brightness')
    tr.optimize_forced_photometry()
    assert abs(star.getBrightness().getValue()) < 1e-3

    tim1.data = mods[0]
    tim2.data = mods[1]
    tr.optimize_forced_photometry()
    assert abs(star.getBrightness().getValue() - trueflux) < 1e-3

    star.getBrightness().setValue(0.)
    np.random.seed(42)
    tim1.data = mods[0] + np.random.normal(size=tim1.shape)
    tim2.data = mods[1] + np.random.normal(size=tim2.shape)
    tr.optimize_forced_photometry()
    assert abs(star.getBrightness().getValue() - trueflux) < 5.


class ConstantSource(MultiParams, T.BasicSource):
    """A source that is 1 on every pixel (test/test_forced_phot.py:11-47)."""
    def __init__(self, br):
        super().__init__(br)

    def getBrightness(self):
        return self.brightness

    def getBrightnesses(self):
        return [self.brightness]

    def getUnitFluxModelPatches(self, *args, **kwargs):
        return [self.getUnitFluxModelPatch(*args, **kwargs)]

    def getUnitFluxModelPatch(self, img, modelMask=None, **kwargs):
        if modelMask is None:
            return Patch(0, 0, np.ones(img.shape, np.float32))
        return Patch(0, 0, np.ones(modelMask.shape, np.float32))

    def getModelPatch(self, img, modelMask=None, **kwargs):
        counts = img.getPhotoCal().brightnessToCounts(self.brightness)
        if counts == 0:
            return None
        return self.getUnitFluxModelPatch(img, modelMask=modelMask) * counts

    @staticmethod
    def getNamedParams():
        return dict(brightness=0)


class ConstantSourceWithPrior(GaussianPriorsMixin, ConstantSource):
    priorStd = 0.
    priors = None

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        if self.priors is None:
            p = _GaussianPriors(None)
            p.add('brightness', 0., self.priorStd, param=ConstantSource(T.Flux(0.)))
            self.__class__.priors = p
        self.gpriors = self.priors


def _pixel_tim(val):
    return T.Image(data=np.zeros((1, 1), np.float32) + val, inverr=np.ones((1, 1), np.float32),
                   photocal=T.LinearPhotoCal(1.))


def test_forced_phot_pixel():
    """test/test_forced_phot.py:69-86."""
    tim = _pixel_tim(42.)
    src = ConstantSource(T.Flux(10.))
    tr = T.Tractor([tim], [src])
    tr.freezeParam('images')
    tr.optimize_forced_photometry()
    assert abs(src.getParams()[0] - 42.) < 1e-6


def test_forced_phot_pixel_with_prior():
    """test/test_forced_phot.py:88-140: the flux prior N(0, 1) halves the
    estimate; with inverr 2 the estimate is the inverse-variance mean."""
    val = 4.
    tim = _pixel_tim(val)
    src = ConstantSource(T.Flux(0.))
    tr = T.Tractor([tim], [src])
    tr.freezeParam('images')
    tr.optimize_forced_photometry()
    assert abs(src.getParams()[0] - val) < 1e-6

    ConstantSourceWithPrior.priorStd = 1.
    src = ConstantSourceWithPrior(T.Flux(0.))
    tr = T.Tractor([tim], [src])
    tr.freezeParam('images')
    tr.optimize_forced_photometry(priors=True)
    assert abs(src.getParams()[0] - val / 2.) < 1e-6
    src.setParams([1.])
    tr.optimize_forced_photometry(priors=True)
    assert abs(src.getParams()[0] - val / 2.) < 1e-6

    tim.inverr[:, :] = 2.
    src.setParams([0.])
    tr.optimize_forced_photometry(priors=True)
    pval = (4. * val + 1. * 0.) / (4. + 1.)
    assert abs(src.getParams()[0] - pval) < 1e-6


def test_optimize_loop_recovers_galaxy():
    """Tractor.optimize_loop with the constrained (GPU plugin) optimizer on a
    noiseless exponential galaxy: flux and shape return to the truth.  The
    position is frozen: with the default NullWCS the reference's position
    derivatives carry the cd^-1 = 3600 factor (wcs.py:27-43, ducks.py:413-423),
    so position steps are 3600x too small and positions barely move."""
    psf = T.GaussianMixturePSF(1., 0., 0., 4., 4., 0.)
    H = W = 64
    truth = T.ExpGalaxy(T.PixPos(31.7, 32.4), T.Flux(500.), T.EllipseESoft(np.log(3.), 0.2, -0.1))
    tim = T.Image(data=np.zeros((H, W), np.float32), inverr=np.ones((H, W), np.float32) * 10.,
                  psf=psf, photocal=T.LinearPhotoCal(1.))
    tr0 = T.Tractor([tim], [truth])
    tim.data = tr0.getModelImage(0).astype(np.float32)
    gal = T.ExpGalaxy(T.PixPos(31., 33.), T.Flux(400.), T.EllipseESoft(np.log(2.5), 0., 0.))
    gal.pos = T.PixPos(31.7, 32.4)
    gal.freezeParam('pos')
    tr = T.Tractor([tim], [gal], optimizer=T.GpuOptimizer())
    tr.freezeParam('images')
    lnp0 = tr.getLogProb()
    r = tr.optimize_loop()
    assert r['steps'] < 100
    assert tr.getLogProb() > lnp0
    np.testing.assert_allclose(gal.getParams(), truth.getParams()[2:], atol=2e-3, rtol=2e-3)


def test_lsqr_update_matches_dense_lstsq():
    """One LsqrOptimizer update direction against numpy's lstsq on the same
    column-scaled derivative matrix (FP32 normal equations: 1e-4 relative)."""
    psf = T.GaussianMixturePSF(1., 0., 0., 4., 4., 0.)
    H = W = 48
    rng = np.random.RandomState(3)
    tim = T.Image(data=rng.normal(size=(H, W)).astype(np.float32),
                  inverr=np.ones((H, W), np.float32), psf=psf, photocal=T.LinearPhotoCal(1.))
    src = T.PointSource(T.PixPos(23.3, 24.6), T.Flux(300.))
    tr = T.Tractor([tim], [src])
    tr.freezeParam('images')
    opt = T.LsqrOptimizer()
    X = opt.getLinearUpdateDirection(tr, priors=False)
    derivs = tr.getDerivs()
    cols = []
    for param in derivs:
        c = np.zeros(H * W)
        for d, img in param:
            m = np.zeros((H, W))
            d.addTo(m)
            c += m.ravel()
        cols.append(c)
    A = np.array(cols).T
    sc = np.sqrt((A ** 2).sum(0))
    b = tr.getChiImage(0).ravel()
    Xs = np.linalg.lstsq(A / sc, b, rcond=None)[0] / sc
    np.testing.assert_allclose(X, Xs, rtol=1e-3, atol=1e-4 * np.abs(Xs).max())


def test_forced_phot_fitstats_and_variance():
    """optimize_forced_photometry(fitstats=True, variance=True) (the FitStats
    of optimize.py:233-376 and _get_iv, 378-423, on the device) against the
    oracle restatement over the optimizer's own unit-flux models: two
    images, one with a constant sky and a masked block, two overlapping point
    sources and an isolated one."""
    import oracle as O
    W, H = 80, 60
    rng = np.random.default_rng(8)
    tims = []
    for k, (sig, skyv) in enumerate(((2., 0.), (3., 1.5))):
        ie = np.ones((H, W))
        if k == 1:
            ie[20:26, 30:41] = 0.
        tims.append(T.Image(data=np.zeros((H, W)), inverr=ie,
                            psf=T.NCircularGaussianPSF([sig], [1.]), photocal=T.LinearPhotoCal(2.),
                            sky=T.ConstantSky(skyv)))
    srcs = [T.PointSource(T.PixPos(30., 25.), T.Flux(40.)),
            T.PointSource(T.PixPos(34., 27.), T.Flux(25.)),
            T.PointSource(T.PixPos(65., 45.), T.Flux(60.))]
    tr = T.Tractor(tims, srcs)
    for tim, mod in zip(tims, tr.getModelImages()):
        tim.data = mod + rng.normal(size=mod.shape)
    for s in srcs:
        s.getBrightness().setValue(10.)
        s.freezeAllBut('brightness')
    tr.freezeParam('images')
    res = tr.optimize_forced_photometry(fitstats=True, variance=True)
    fs = res.fitstats
    opt = tr.optimizer
    umodels = opt._get_umodels(tr, srcs, tims, 0.)
    fluxes = np.array(tr.catalog.getParams())
    per = []
    for tim, umods in zip(tims, umodels):
        scale = tim.getPhotoCal().getScale()
        pts = []
        for um in umods:
            p = np.asarray(um.patch, np.float64)
            X0, Y0 = max(um.x0, 0), max(um.y0, 0)
            X1, Y1 = min(um.x0 + p.shape[1], W), min(um.y0 + p.shape[0], H)
            pts.append((X0, Y0, p[Y0 - um.y0:Y1 - um.y0, X0 - um.x0:X1 - um.x0]))
        skyv = tim.getSky().val
        mod = np.asarray(tr.getModelImage(tim), np.float64) - skyv
        per.append(O.forced_fitstats(pts, fluxes * scale, mod, tim.getImage(), tim.getInvError(),
                                     scale=scale, sky=skyv))
    tot = lambda k: sum(p[k] for p in per)
    np.testing.assert_allclose(res.IV, tot('iv'), rtol=1e-5)
    for k in ('prochi2', 'pronpix', 'promasked', 'proflux'):
        np.testing.assert_allclose(getattr(fs, k), tot(k), rtol=1e-5, atol=1e-9, err_msg=k)
    np.testing.assert_array_equal(fs.npix, tot('npix'))
    np.testing.assert_allclose(fs.profracflux, tot('ffnum') / np.maximum(1., tot('ffden')),
                               rtol=1e-5, atol=1e-9)
    np.testing.assert_allclose(fs.fracin, tot('fracin_num') / np.maximum(1., tot('counted')),
                               rtol=1e-5)
    np.testing.assert_allclose(fs.imchisq, [p['imchisq'] for p in per], rtol=1e-5)
    np.testing.assert_array_equal(fs.imnpix, [p['imnpix'] for p in per])
    np.testing.assert_allclose(fs.sky, [0., 1.5])
    # the overlapping pair sees each other's flux, the isolated source none
    assert fs.profracflux[0] > 0.05 and fs.profracflux[1] > 0.05
    assert fs.profracflux[2] < 1e-6


@pytest.mark.parametrize('n', [0.5, 2.5, 5.5])
def test_sersic_patch_matches_oracle(n):
    """SersicGalaxy (sersic.py:362-500) through the device FFT render against
    the FP64 oracle render of the same sheared mixture (SersicMixture at n),
    and its index derivative against the oracle's forward difference."""
    from oracle import tractor_oracle as O
    psf0 = T.GaussianMixturePSF(1., 0., 0., 6., 6., -1.)
    psfpatch = psf0.getPointSourcePatch(24., 24., modelMask=ModelMask(0, 0, 50, 50))
    pix = psfpatch.patch[:49, :49].astype(np.float32)
    tim = _tim(T.PixelizedPSF(pix))
    shape = T.EllipseESoft(np.log(2.2), 0.15, -0.1)
    gal = T.SersicGalaxy(T.PixPos(49.3, 50.2), T.Flux(100.), shape, T.SersicIndex(n))
    mm = ModelMask(25, 25, 50, 50)
    p = gal.getModelPatch(tim, modelMask=mm)

    def oracle_patch(nn):
        mix = T.SersicMixture.getProfile(nn)
        var = O.shear_mixture(mix.var, np.eye(2) * 3600.,
                              O.ellipse_esoft_basis(np.log(2.2), 0.15, -0.1))
        return O.render_galaxy_fft(100. * mix.amp, var, 49.3, 50.2, pix, mm=(25, 25, 50, 50))
    x0, y0, ref = oracle_patch(n)
    assert (p.x0, p.y0) == (x0, y0)
    print('  sersic patch n %.2f max err / peak %.3e' % (n, np.abs(p.patch - ref).max() /
                                                         np.abs(ref).max()), flush=True)
    np.testing.assert_allclose(p.patch, ref, rtol=0, atol=1e-5 * np.abs(ref).max())
    gal.freezeAllBut('sersicindex')
    d = gal.getParamDerivatives(tim, modelMask=mm)
    assert len(d) == 1
    _, _, ref1 = oracle_patch(n + 0.01)
    dref = (ref1 - ref) / 0.01
    np.testing.assert_allclose(d[0].patch, dref, rtol=0, atol=5e-3 * np.abs(dref).max())


def test_optimize_loop_recovers_sersic_index():
    """optimize_loop on a noiseless Sersic galaxy (n = 2.5) from n = 2.0 and a
    perturbed shape and flux: every free parameter returns to the truth."""
    psf = T.GaussianMixturePSF(1., 0., 0., 4., 4., 0.)
    H = W = 64
    truth = T.SersicGalaxy(T.PixPos(31.7, 32.4), T.Flux(500.),
                           T.EllipseESoft(np.log(3.), 0.2, -0.1), T.SersicIndex(2.5))
    tim = T.Image(data=np.zeros((H, W), np.float32), inverr=np.ones((H, W), np.float32) * 10.,
                  psf=psf, photocal=T.LinearPhotoCal(1.))
    tim.data = T.Tractor([tim], [truth]).getModelImage(0).astype(np.float32)
    gal = T.SersicGalaxy(T.PixPos(31.7, 32.4), T.Flux(420.), T.EllipseESoft(np.log(2.6), 0.1, 0.),
                         T.SersicIndex(2.0))
    gal.freezeParam('pos')
    tr = T.Tractor([tim], [gal], optimizer=T.GpuOptimizer())
    tr.freezeParam('images')
    lnp0 = tr.getLogProb()
    r = tr.optimize_loop()
    assert r['steps'] < 100 and tr.getLogProb() > lnp0
    np.testing.assert_allclose(gal.getParams(), truth.getParams()[2:], atol=5e-3, rtol=5e-3)


def _scene_catalog(n, psf_stamp, rng):
    srcs = []
    for i in range(n):
        x, y = rng.uniform(10, 90, 2)
        k = i % 3
        if k == 0:
            srcs.append(T.PointSource(T.PixPos(x, y), T.Flux(rng.uniform(50, 500))))
        else:
            cls = T.ExpGalaxy if k == 1 else T.DevGalaxy
            srcs.append(cls(T.PixPos(x, y), T.Flux(rng.uniform(50, 500)),
                            T.EllipseESoft(np.log(rng.uniform(1., 4.)), rng.uniform(-0.3, 0.3),
                                           rng.uniform(-0.3, 0.3))))
    return srcs


def test_model_image_and_derivs_are_batched():
    """getModelImage and getDerivs plan every source into one RenderPlan:
    one synchronisation and one render call per bucket (FFT size, stamp),
    not one per source; the pixels equal the per-source renders bit for bit
    (each render is one workgroup's, whatever else shares its launch)."""
    from tractor_amd.plan import RenderPlan
    import oracle.tractor_oracle as O
    from tractor_amd import workloads as W
    rng = np.random.default_rng(20261206)
    psf = T.PixelizedPSF(W.moffat_stamp(31, 4.0, 3.5))
    tim = _tim(psf)
    srcs = _scene_catalog(24, psf, rng)
    tr = T.Tractor([tim], srcs)
    st = RenderPlan.stats
    r0, c0 = st['runs'], st['calls']
    mod = tr.getModelImage(0)
    runs, calls = st['runs'] - r0, st['calls'] - c0
    assert runs == 1
    # point sources: one shift call (one stamp shape); galaxies: one call
    # per FFT size in use
    assert calls <= 1 + 3, calls
    ref = np.zeros_like(mod)
    for s in srcs:
        p = s.getModelPatch(tim)
        if p is not None:
            p.addTo(ref)
    assert np.array_equal(mod, ref)
    # derivatives: every source's renders (unit + shape steps) in one plan
    tr.freezeParam('images')
    r0, c0 = st['runs'], st['calls']
    allderivs = tr.getDerivs()
    assert st['runs'] - r0 == 1 and st['calls'] - c0 <= 1 + 3
    k = 0
    for s in srcs:
        own = s.getParamDerivatives(tim)
        for d in own:
            got = allderivs[k]
            k += 1
            if d is None:
                assert got == []
                continue
            (gd, gimg), = got
            assert gimg is tim and (gd.x0, gd.y0) == (d.x0, d.y0)
            assert np.array_equal(gd.patch, d.patch)
    assert k == len(allderivs)


def test_line_search_is_one_plan():
    """A constrained line search renders all of its trial steps in one plan
    and picks the step the sequential rule picks."""
    from tractor_amd.plan import RenderPlan
    from tractor_amd import workloads as W
    rng = np.random.default_rng(20261207)
    psf = T.PixelizedPSF(W.moffat_stamp(31, 4.0, 3.5))
    truth = _scene_catalog(6, psf, rng)
    tim = _tim(psf)
    tim.data = T.Tractor([tim], truth).getModelImage(0) + rng.normal(size=tim.shape).astype(
        np.float32)
    start = [s.copy() for s in truth]
    for s in start:
        s.brightness.setParams([s.brightness.getParams()[0] * 1.2])
    tr = T.Tractor([tim], start, optimizer=T.ConstrainedOptimizer())
    tr.freezeParam('images')
    opt = tr.optimizer
    X = opt.getLinearUpdateDirection(tr)
    p0 = np.array(tr.getParams())
    steps = opt.getParameterSteps(tr, X, None)
    # the sequential reference rule on the same steps, one getLogProb each
    lnp0 = tr.getLogProb()
    best, ib = lnp0, None
    for i, (a, p, _, _) in enumerate(steps):
        tr.setParams(p)
        lnp = tr.getLogProb()
        if not np.isfinite(lnp) or lnp < best - 1.:
            break
        if lnp > best:
            best, ib = lnp, i
    tr.setParams(p0)
    st = RenderPlan.stats
    r0 = st['runs']
    dlnp, alpha = opt.tryUpdates(tr, X)
    assert st['runs'] - r0 == 1
    assert ib is not None and alpha == steps[ib][0]
    assert abs(dlnp - (best - lnp0)) <= 1e-9 * abs(best)
    assert np.allclose(tr.getParams(), steps[ib][1], rtol=0, atol=0)
