"""Edge cases on the device: empty batches, the largest FFT size, windows
that miss the rendered support, one-pixel windows, sources outside the
image, and a batch of one cutout through the CNN."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return 'cuda'


def _sheared(kind, shape):
    import oracle as O
    amp, var = O.exp_mixture() if kind == 1 else O.dev_mixture()
    return amp, O.shear_mixture(var, np.eye(2) * 3600., O.ellipse_esoft_basis(*shape))


def test_empty_batches_are_no_ops(dev):
    import torch
    from tractor_amd import _lib
    lib, h = _lib.load(), _lib.handle(0)
    out = torch.full((16,), 7.0, device=dev)
    assert lib.trh_render_fft_batch(h.value, 64, None, 0, _lib.ptr(out), _lib.ptr(out), None, None,
                                    None, None) == 0
    assert lib.trh_fp_rhs(h.value, 0, None, None, None, None, None, None, 4, None, None, None) == 0
    assert lib.trh_fp_pair_dots(h.value, None, None, 0, None, None, None, None, 4, None, None) == 0
    torch.cuda.synchronize()
    assert torch.all(out == 7.0)


def test_largest_fft_size_matches_oracle(dev):
    """A 260x260 modelMask needs sz = 512 (halfsize 130, galaxy.py:337-351)."""
    import torch
    import oracle as O
    from tractor_amd import render as R, workloads as W
    from tractor_amd._lib import FFT_JOB, TRH_STORE
    print('sz 512 render', flush=True)
    psf = W.moffat_stamp(63, 4.0, 3.5)
    sp = R.PsfSpectra(psf, dev)
    amp, var = _sheared(2, O.esoft_from_e(6., 0.3, 0.1))
    P = 260
    sz, J, win = R.galaxy_fft_job(amp, var, 130.3, 129.6, psf.shape, mm=(0, 0, P, P))
    assert sz == 512
    J['out_offset'], J['out_stride'], J['mode'] = 0, P, TRH_STORE
    out = torch.zeros(P * P, device=dev)
    R.render_fft_jobs(sz, np.array([J], FFT_JOB), sp, out=out, device=dev)
    torch.cuda.synchronize()
    _, _, G = O.render_galaxy_fft(amp, var, 130.3, 129.6, psf, mm=(0, 0, P, P))
    got = out.cpu().numpy().reshape(P, P)
    e = np.max(np.abs(got - G)) / np.max(np.abs(G))
    print('  sz 512 max err / peak %.3e' % e, flush=True)
    assert e <= 1e-5


def test_thin_and_one_pixel_windows(dev):
    """modelMasks of 1x1, 3x40 and 40x8 pixels that contain the source
    (the reference refuses a source outside a non-hybrid mask,
    galaxy.py:354-374): the halfsize rule still sizes the FFT from the
    mask edges."""
    import torch
    import oracle as O
    from tractor_amd import render as R, workloads as W
    from tractor_amd._lib import FFT_JOB, TRH_STORE
    psf = W.moffat_stamp(31, 3.0, 3.5)
    sp = R.PsfSpectra(psf, dev)
    amp, var = _sheared(1, O.esoft_from_e(2., 0.1, 0.0))
    for px, py, mm in [(12.0, 13.0, (12, 13, 1, 1)), (12.4, 13.7, (11, 0, 3, 40)),
                       (12.4, 13.7, (12, 13, 40, 8)), (12.4, 13.7, (-20, -30, 70, 2 + 30 + 13))]:
        print('window', px, py, mm, flush=True)
        sz, J, win = R.galaxy_fft_job(amp, var, px, py, psf.shape, mm=mm)
        J['out_offset'], J['out_stride'], J['mode'] = 0, mm[2], TRH_STORE
        out = torch.full((mm[2] * mm[3],), 123.0, device=dev)
        R.render_fft_jobs(sz, np.array([J], FFT_JOB), sp, out=out, device=dev)
        torch.cuda.synchronize()
        _, _, G = O.render_galaxy_fft(amp, var, px, py, psf, mm=mm)
        got = out.cpu().numpy().reshape(mm[3], mm[2])
        print('  max err %.3e peak %.3e' % (np.max(np.abs(got - G)), np.max(np.abs(G))), flush=True)
        assert np.max(np.abs(got - G)) <= 1e-5 * max(np.max(np.abs(G)), 1e-30) + 1e-12, mm


def test_forced_photometry_sources_outside_the_image(dev):
    """Sources whose patches miss the image get flux 0 and iv 0 (LSQR leaves
    empty columns at zero); the others are unaffected."""
    import torch
    from tractor_amd import forced as F, workloads as W
    size = 96
    kind = np.array([0, 1, 0, 2, 0], np.int32)
    pos = np.array([[40., 40.], [60., 50.], [-80., 30.], [300., 300.], [20., 70.]])
    shape = np.array([[0, 0, 0], [0.7, 0.1, 0.], [0, 0, 0], [0.5, 0., 0.1], [0, 0, 0]], np.float64)
    basis = W.psf_grid_stamps()
    idx, w = W.bilinear_psf_weights(np.clip(pos[:, 0], 0, size - 1), np.clip(pos[:, 1], 0, size - 1),
                                    size, size)
    rng = np.random.default_rng(1)
    data = torch.from_numpy(rng.normal(size=(size, size)).astype(np.float32)).cuda()
    inv = torch.ones_like(data)
    fp = F.BatchForcedPhotometry(data, inv, basis, idx, w, kind, pos, shape, tile=32)
    f = fp.run()
    assert np.all(fp.info == 0)
    assert f[2] == 0. and f[3] == 0. and fp.iv[2] == 0. and fp.iv[3] == 0.
    assert np.all(fp.iv[[0, 1, 4]] > 0) and np.all(np.isfinite(f))


def test_cnn_single_cutout(dev):
    import torch
    from tractor_amd import surrogate as S
    net = S.SurrogateCNN(max_batch=1)
    x = torch.rand((1, 32, 32, 1), device=dev).to(torch.bfloat16)
    out = net.predict(x, 1)
    torch.cuda.synchronize()
    assert out.shape == (1, 9) and torch.all(torch.isfinite(out))
    with pytest.raises(ValueError):
        net.forward(torch.rand((2, 32, 32, 1), device=dev).to(torch.bfloat16), 2)


def _interior(sz, J):
    """The render kernel filters the spectrum (Lanczos as a frequency
    response) when every tap of every output pixel stays inside the box."""
    sy, sx, mh, mw = int(J['sy']), int(J['sx']), int(J['mh']), int(J['mw'])
    ylo, yhi = max(0, -sy), min(sz, mh - sy)
    return ylo >= 3 and yhi <= sz - 3 and -sx >= 3 and mw - sx <= sz - 3


@pytest.mark.parametrize('sz_target', [64, 128, 256])
def test_spectrum_filter_and_spatial_paths(dev, sz_target):
    """Windows on both sides of the interior rule, for one FFT size: the
    spectrum-filter path and the spatial Lanczos path agree with the oracle."""
    import torch
    import oracle as O
    from tractor_amd import render as R, workloads as W
    from tractor_amd._lib import FFT_JOB, TRH_STORE
    psf = W.moffat_stamp(31, 3.0, 3.5)
    sp = R.PsfSpectra(psf, dev)
    amp, var = _sheared(2, O.esoft_from_e(3., 0.3, -0.2))
    rng = np.random.default_rng(sz_target)
    seen = set()
    Ps = list(range(sz_target // 2 - 1, sz_target - 1, max(1, sz_target // 16))) + \
        [sz_target - 3, sz_target - 2]
    for P in Ps:
        for _ in range(2):
            px, py = P / 2. + rng.uniform(-0.5, 0.5), P / 2. + rng.uniform(-0.5, 0.5)
            sz, J, win = R.galaxy_fft_job(amp, var, px, py, psf.shape, mm=(0, 0, P, P))
            if sz != sz_target:
                continue
            path = _interior(sz, J)
            seen.add(path)
            J['out_offset'], J['out_stride'], J['mode'] = 0, P, TRH_STORE
            out = torch.zeros(P * P, device=dev)
            R.render_fft_jobs(sz, np.array([J], FFT_JOB), sp, out=out, device=dev)
            torch.cuda.synchronize()
            _, _, G = O.render_galaxy_fft(amp, var, px, py, psf, mm=(0, 0, P, P))
            got = out.cpu().numpy().reshape(P, P)
            assert np.max(np.abs(got - G)) <= 1e-5 * np.max(np.abs(G)), (sz, P, px, py, path)
    assert seen == {True, False}, seen
