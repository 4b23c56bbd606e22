"""The device EM fit of a mixture of Gaussians to PSF stamps
(trh_em_fit_psf; GaussianMixturePSF.fromStamp, psf.py:537-597, and
HybridPixelizedPSF(gauss=None), psf.py:600-625) against the FP64 oracle
restatement of em_fit_2d_reg (emfit.i:274).  Both run the same EM steps in
FP64; they differ only in summation order, so the fitted parameters agree
to 1e-8 relative."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module', autouse=True)
def _need_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')


def _oracle_fit(stamp, K, steps, alpha=0.):
    import oracle as O
    st = np.asarray(stamp, np.float64)
    st = np.maximum(st / st.sum(), 0.)
    w, mu, var = O.em_init_params(K)
    NY, NX = st.shape
    return O.em_fit_2d_reg(st, -(NX // 2), -(NY // 2), w, mu, var, alpha, steps)


@pytest.mark.parametrize('K,steps,alpha', [(1, 50, 0.), (2, 1000, 0.), (3, 300, 0.5)])
def test_em_fit_matches_oracle(K, steps, alpha):
    from tractor_amd import psf as P, workloads as W
    stamps = [W.moffat_stamp(63, 4.0, 3.5)] + list(W.psf_grid_stamps()[[0, 4, 8]])
    print('em fit K=%d steps=%d on %d stamps' % (K, steps, len(stamps)), flush=True)
    shapes = {s.shape for s in stamps}
    for shp in shapes:
        group = [s for s in stamps if s.shape == shp]
        got = P.fit_mog_to_stamps(group, N=K, alpha=alpha, emsteps=steps)
        for s, (a, m, v) in zip(group, got):
            ra, rm, rv = _oracle_fit(s, K, steps, alpha)
            np.testing.assert_allclose(a, ra, rtol=1e-8, atol=1e-12)
            np.testing.assert_allclose(m, rm, rtol=1e-8, atol=1e-9)
            np.testing.assert_allclose(v, rv, rtol=1e-8, atol=1e-9)


def test_hybrid_psf_fits_its_mixture():
    """HybridPixelizedPSF(pix) with no mixture fits N = 2 components to the
    stamp (the reference's default), matching the oracle; the mixture
    carries the stamp's flux (amplitudes sum to ~1)."""
    import tractor_amd as T
    from tractor_amd import workloads as W
    stamp = W.moffat_stamp(63, 4.0, 3.5)
    pix = T.PixelizedPSF(stamp)
    hyb = T.HybridPixelizedPSF(pix)
    mog = hyb.getMixtureOfGaussians()
    ra, rm, rv = _oracle_fit(pix.getImage(0., 0.), 2, 1000)
    np.testing.assert_allclose(mog.amp, ra, rtol=1e-8)
    np.testing.assert_allclose(mog.mean, rm, rtol=1e-8, atol=1e-9)
    np.testing.assert_allclose(mog.var, rv, rtol=1e-8, atol=1e-9)
    assert 0.9 < np.sum(mog.amp) <= 1.0 + 1e-12
