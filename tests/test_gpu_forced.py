"""Batched forced photometry on the device (tractor_amd.forced) against the
FP64 oracle (oracle.forced_unit_patch + forced_phot_lsq) on a 384x384 scene
of 240 overlapping sources with a bilinearly interpolated 3x3 PSF grid.
Unit-flux patches agree to FP32 rendering accuracy (2e-5 of each patch's
peak); fluxes within 1e-4 relative or 0.01 sigma (sigma = iv^-1/2)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module', autouse=True)
def _need_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')


def test_forced_photometry_matches_oracle():
    import torch
    import oracle as O
    from tractor_amd import forced as F, workloads as W
    size, n = 384, 240
    rng = np.random.default_rng(31)
    kind, pos, flux, shape = W.draw_forced_scene(n, size, rng=rng)
    basis = W.psf_grid_stamps()
    idx, w = W.bilinear_psf_weights(pos[:, 0], pos[:, 1], size, size)
    stamps = np.einsum('nq,nqij->nij', w.astype(np.float64), basis[idx].astype(np.float64))
    print('oracle patches', flush=True)
    patches = [O.forced_unit_patch(int(kind[s]), pos[s, 0], pos[s, 1], shape[s], stamps[s],
                                   (size, size)) for s in range(n)]
    model = np.zeros((size, size))
    for pt, f in zip(patches, flux):
        if pt is not None:
            x0, y0, p = pt
            model[y0:y0 + p.shape[0], x0:x0 + p.shape[1]] += f * p
    data = (model + rng.normal(size=model.shape)).astype(np.float32)
    inverr = np.ones((size, size), np.float32)
    ref, iv_ref = O.forced_phot_lsq(patches, data.astype(np.float64), inverr.astype(np.float64))
    print('device forced photometry', flush=True)
    fp = F.BatchForcedPhotometry(torch.from_numpy(data).cuda(), torch.from_numpy(inverr).cuda(),
                                 basis, idx, w, kind, pos, shape, tile=128)
    got = fp.run(tol=1e-13)
    print('pcg iterations', fp.iterations, 'resid', fp.resid, 'timing', fp.timing, flush=True)
    assert np.all(fp.info == 0)
    # unit-flux patches
    buf = fp.patches.cpu().numpy()
    inv = np.empty(n, np.int64)
    inv[fp.order] = np.arange(n)
    for s in range(n):
        i = inv[s]
        x0, y0, ww, hh = fp.win[i]
        if patches[s] is None:
            assert ww * hh == 0
            continue
        rx0, ry0, p = patches[s]
        assert (x0, y0, ww, hh) == (rx0, ry0, p.shape[1], p.shape[0])
        g = buf[fp.poff[i]:fp.poff[i] + ww * hh].reshape(hh, ww)
        assert np.max(np.abs(g - p)) <= 2e-5 * np.max(np.abs(p)) + 1e-12, s
    np.testing.assert_allclose(fp.iv, iv_ref, rtol=1e-4, atol=1e-12)
    sig = 1. / np.sqrt(np.maximum(iv_ref, 1e-300))
    err = np.abs(got - ref)
    ok = (err <= 1e-4 * np.abs(ref)) | (err <= 0.01 * sig)
    print('flux parity: %d/%d; max err/sigma %.3g' % (ok.sum(), n, np.max(err / sig)), flush=True)
    assert ok.all()


def test_device_overlap_pattern_matches_host_rule():
    """trh_fp_overlaps (cells of tile x tile pixels, a per-source wavefront
    scanning the cells within reach) gives exactly the CSR pattern of the
    host's cell-list rule (overlap_pairs + csr_of_pairs), sorted columns and
    the diagonal's positions included."""
    import torch
    from tractor_amd import forced as F, workloads as W
    size, n = 512, 1500
    rng = np.random.default_rng(5)
    kind, pos, flux, shape = W.draw_forced_scene(n, size, rng=rng)
    pos[:7] = [[-40., 10.], [600., 30.], [5., -3.], [511.5, 511.5], [0., 0.], [250., 700.], [-200., -200.]]
    basis = W.psf_grid_stamps()
    idx, w = W.bilinear_psf_weights(np.clip(pos[:, 0], 0, size - 1), np.clip(pos[:, 1], 0, size - 1),
                                    size, size)
    data = torch.zeros((size, size), dtype=torch.float32, device='cuda')
    fp = F.BatchForcedPhotometry(data, torch.ones_like(data), basis, idx, w, kind, pos, shape, tile=64)
    fp.plan()
    torch.cuda.synchronize()
    rp, col, _ = F.csr_of_pairs(F.overlap_pairs(fp.win), n)
    assert np.array_equal(fp.row_ptr, rp)
    dcol = fp.d_col.cpu().numpy()
    assert np.array_equal(dcol, col)
    diag = fp.d_diag.cpu().numpy()
    assert np.array_equal(dcol[diag], np.arange(n))
    assert fp.reach >= 1
