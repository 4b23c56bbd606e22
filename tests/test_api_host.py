"""Host-side logic of the drop-in API (no device calls): parameter plumbing
(test/test_tractor.py:13-51 and test/test_galaxy.py:45-57 with PixPos),
patches and model masks, the constrained step generator, Gaussian priors,
and that the product path refuses to run without the HIP library's device."""
import numpy as np
import pytest

import tractor_amd as T
from tractor_amd.patch import Patch, ModelMask
from tractor_amd.utils import _GaussianPriors, GaussianPriorsMixin


def test_expgal_params():
    pos = T.PixPos(12., 45.)
    shape = T.GalaxyShape(7, 0.8, 9)
    gal = T.ExpGalaxy(pos, T.NanoMaggies(r=6.), shape)
    assert str(gal) == ('ExpGalaxy at pixel (12.00, 45.00) with NanoMaggies: r=20.555 and '
                        'Galaxy Shape: re=7.00, ab=0.80, phi=9.0')
    assert gal.getParams() == [12., 45., 6., 7, 0.8, 9]
    gal.shape.re = 7.7
    assert shape.re == 7.7
    assert gal.getParams() == [12., 45., 6., 7.7, 0.8, 9]
    gal.shape = T.GalaxyShape(7.77, 0.8, 9)
    assert gal.shape.re == 7.77 and shape.re == 7.7


def test_galaxy_strings():
    """The reference formats NanoMaggies magnitudes with %.3f
    (tractor/brightness.py:173-184)."""
    gal = T.ExpGalaxy(T.PixPos(49.5, 50.), T.NanoMaggies(g=1., r=2.), T.GalaxyShape(2., 0.5, 45.))
    assert str(gal) == ('ExpGalaxy at pixel (49.50, 50.00) with NanoMaggies: g=22.500, r=21.747 '
                        'and Galaxy Shape: re=2.00, ab=0.50, phi=45.0')
    assert repr(gal).startswith('ExpGalaxy(pos=PixPos[49.5, 50.0], brightness=NanoMaggies')
    assert gal.getParamNames() == ['pos.x', 'pos.y', 'brightness.g', 'brightness.r', 'shape.re',
                                   'shape.ab', 'shape.phi']


def test_tractor_freeze_thaw():
    tim = T.Image(data=np.zeros((10, 10), np.float32), inverr=np.ones((10, 10), np.float32),
                  psf=T.GaussianMixturePSF(1., 0., 0., 4., 4., 0.), photocal=T.LinearPhotoCal(1.))
    s1 = T.PointSource(T.PixPos(3., 4.), T.Flux(5.))
    s2 = T.ExpGalaxy(T.PixPos(6., 7.), T.Flux(8.), T.EllipseESoft(0.1, 0.2, 0.3))
    tr = T.Tractor([tim], [s1, s2])
    tr.freezeParam('images')
    assert tr.numberOfParams() == 3 + 6
    assert tr.getParams() == [3., 4., 5., 6., 7., 8., 0.1, 0.2, 0.3]
    s2.freezeParam('shape')
    assert tr.numberOfParams() == 6
    tr.setParams([1., 2., 3., 4., 5., 6.])
    assert s1.getParams() == [1., 2., 3.] and s2.getAllParams() == [4., 5., 6., 0.1, 0.2, 0.3]
    s1.freezeAllBut('brightness')
    assert tr.getParamNames() == ['catalog.source0.brightness.Flux', 'catalog.source1.pos.x',
                                  'catalog.source1.pos.y', 'catalog.source1.brightness.Flux']
    assert len(list(tr.catalog.getThawedSources())) == 2
    # frozen sources are those frozen in the catalog (utils.py:1030-1040)
    tr.catalog.freezeParam(1)
    assert list(tr.catalog.getFrozenSources()) == [s2]
    assert list(tr.catalog.getThawedSources()) == [s1]


def test_ellipse_esoft_bounds_and_steps():
    e = T.EllipseESoft(np.log(2.), 0.3, -0.2)
    assert abs(e.re - 2.) < 1e-12
    assert e.getMaxStep()[0] is not None
    src = T.ExpGalaxy(T.PixPos(1., 2.), T.Flux(3.), e)
    assert len(src.getLowerBounds()) == src.numberOfParams() == len(src.getMaxStep())


def test_patch_arithmetic():
    p = Patch(2, 3, np.arange(12, dtype=np.float32).reshape(3, 4))
    img = np.zeros((5, 5), np.float32)
    p.addTo(img)
    assert img[3, 2] == 0 and img[5 - 1, 5 - 1] == 6.
    q = p.copy()
    q.clipTo(5, 5)
    assert q.shape == (2, 3) and (q.x0, q.y0) == (2, 3)
    assert np.array_equal(q.getPixelIndices(np.zeros((5, 5))), [17, 18, 19, 22, 23, 24])
    r = p * 2.
    assert np.array_equal(r.patch, p.patch * 2)
    s = p + p
    assert np.array_equal(s.patch, r.patch)
    mm = ModelMask(1, 1, 3, 3)
    assert mm.shape == (3, 3) and mm.x1 == 4 and mm.y1 == 4


def test_constrained_parameter_steps():
    """Step limiting and bounds (tractor/constrained_optimizer.py:84-159),
    compared with the oracle's restatement."""
    from oracle import tractor_oracle as O
    from tractor_amd.constrained_optimizer import ConstrainedOptimizer
    src = T.ExpGalaxy(T.PixPos(10., 10.), T.Flux(100.), T.EllipseESoft(0.5, 0.1, 0.1))
    tr = T.Tractor([], [src])
    X = np.array([0.5, -0.3, 40., 2.0, 0.3, -0.2])
    opt = ConstrainedOptimizer()
    steps = opt.getParameterSteps(tr, X, None)
    ref = O.parameter_steps(np.array(src.getParams()), X, tr.getMaxStep(), tr.getLowerBounds(),
                            tr.getUpperBounds())
    assert len(steps) == len(ref)
    for (a, p, sl, hl), (ra, rp, rsl, rhl) in zip(steps, ref):
        assert abs(a - ra) < 1e-15 and sl == rsl and hl == rhl
        np.testing.assert_allclose(p, rp, rtol=0, atol=1e-14)


def test_gaussian_prior_derivatives():
    """A prior N(0.3, 2) on a ParamList element: log prior and its single
    least-squares row, which moves to the liquid index when an earlier
    element is frozen (tractor/utils.py:92-133)."""
    s = T.GalaxyShape(2., 0.5, 45.)
    s.addGaussianPrior('ab', 0.3, 2.)
    assert abs(s.getLogPrior() - (-0.5 * (0.2 / 2.) ** 2)) < 1e-12
    rA, cA, vA, pb, mub = s.getLogPriorDerivatives()
    assert cA == [1] and np.allclose(vA[0], [1. / 2.]) and np.allclose(pb[0], [-0.2 / 2.])
    s.freezeParam('re')
    rA, cA, vA, pb, mub = s.getLogPriorDerivatives()
    assert cA == [0] and np.allclose(pb[0], [-0.2 / 2.])
    s.freezeParam('ab')
    assert s.getLogPriorDerivatives() is None


def test_no_cpu_fallback_without_device():
    """Rendering needs the HIP device; on a host without one the call raises
    instead of computing on the CPU."""
    import torch
    if torch.cuda.is_available():
        pytest.skip('device present')
    tim = T.Image(data=np.zeros((20, 20), np.float32), inverr=np.ones((20, 20), np.float32),
                  psf=T.GaussianMixturePSF(1., 0., 0., 4., 4., 0.), photocal=T.LinearPhotoCal(1.))
    tr = T.Tractor([tim], [T.PointSource(T.PixPos(10., 10.), T.Flux(1.))])
    with pytest.raises(Exception):
        tr.getModelImage(0)


def test_autotune_candidates_and_buckets():
    """The autotuner's search space and its bucket rule (no device)."""
    from tractor_amd import autotune as A
    for sz in A.SIZES:
        c = A.candidates(sz)
        assert c and all(t % 64 == 0 and 64 <= t <= 1024 and r >= 0 for t, r in c)
    # 63x63 PSF: P <= 56 -> 64, 64..120 -> 128, 128 -> 256 (the benchmark buckets)
    assert [A.bucket_of_P(P, (63, 63)) for P in (16, 56, 64, 120, 128)] == [64, 64, 128, 128, 256]
