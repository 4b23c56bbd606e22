"""Batched forced photometry on the device (tractor_amd.forced) against the
FP64 oracle (oracle.forced_unit_patch + forced_phot_lsq) on a 384x384 scene
of 240 overlapping sources with a bilinearly interpolated 3x3 PSF grid.
Unit-flux patches agree to FP32 rendering accuracy (1e-5 of each patch's
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
    worst = 0.
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
        worst = max(worst, np.max(np.abs(g - p)) / np.max(np.abs(p)))
        assert np.max(np.abs(g - p)) <= 1e-5 * np.max(np.abs(p)) + 1e-12, s
    print('  unit patches max err / peak %.3e' % worst, flush=True)
    np.testing.assert_allclose(fp.iv, iv_ref, rtol=1e-4, atol=1e-12)
    sig = 1. / np.sqrt(np.maximum(iv_ref, 1e-300))
    err = np.abs(got - ref)
    ok = (err <= 1e-4 * np.abs(ref)) | (err <= 0.01 * sig)
    print('flux parity: %d/%d; max err/sigma %.3g' % (ok.sum(), n, np.max(err / sig)), flush=True)
    assert ok.all()
    # FitStats (optimize.py:233-376) of the device fit.  (1) The oracle
    # restatement over the device's own patches at the device fluxes: the
    # same pixels, so every statistic agrees to FP32 products / FP64 sums and
    # npix (exact non-zeros) exactly.  (2) Over the oracle's patches, which
    # agree to 2e-5 of their peaks: the profile-weighted sums to ~1e-3 (npix
    # is left out there: FP32 and FP64 FFT renders disagree on exact zeros of
    # far-wing pixels).
    fs = fp.fitstats()
    dev_patches = []
    for s in range(n):
        i = inv[s]
        x0, y0, ww, hh = (int(v) for v in fp.win[i])
        dev_patches.append(None if ww * hh == 0 else
                           (x0, y0, buf[fp.poff[i]:fp.poff[i] + ww * hh].reshape(hh, ww)))

    def model_of(pts):
        m = np.zeros((size, size))
        for pt, f in zip(pts, got):
            if pt is not None:
                x0, y0, p = pt
                m[y0:y0 + p.shape[0], x0:x0 + p.shape[1]] += f * p
        return m
    d64, e64 = data.astype(np.float64), inverr.astype(np.float64)
    ref_d = O.forced_fitstats(dev_patches, got, model_of(dev_patches), d64, e64)
    for key in ('prochi2', 'pronpix', 'proflux', 'fracin', 'profracflux', 'promasked'):
        np.testing.assert_allclose(getattr(fs, key), ref_d[key], rtol=1e-5, atol=2e-5,
                                   err_msg=key)
    np.testing.assert_array_equal(fs.npix, ref_d['npix'])
    assert fs.imchisq[0] == pytest.approx(ref_d['imchisq'], rel=1e-6)
    assert fs.imnpix[0] == ref_d['imnpix']
    ref_fs = O.forced_fitstats(patches, got, model_of(patches), d64, e64)
    # (proflux and profracflux of isolated sources sit at the rounding floor
    # of m - u c in the float32 model image: absolute 2e-5)
    for key, rtol, atol in (('prochi2', 3e-3, 1e-6), ('pronpix', 1e-4, 1e-6),
                            ('proflux', 3e-3, 2e-5), ('fracin', 1e-4, 1e-6),
                            ('profracflux', 3e-3, 2e-5)):
        np.testing.assert_allclose(getattr(fs, key), ref_fs[key], rtol=rtol, atol=atol,
                                   err_msg=key)
    assert fs.imchisq[0] == pytest.approx(ref_fs['imchisq'], rel=1e-4)


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


def test_device_plan_matches_host_plan():
    """trh_fp_keys + trh_fp_plan + trh_fp_jobs_bucketed against the host
    restatement (plan_patches) on a scene with sources off the image and on
    its edges: windows, render buckets, job fields, patch offsets and each
    job's rank inside its bucket."""
    import torch
    from tractor_amd import forced as F, workloads as W
    from tractor_amd._lib import FFT_JOB
    size, n = 512, 1500
    rng = np.random.default_rng(11)
    kind, pos, flux, shape = W.draw_forced_scene(n, size, rng=rng)
    pos[:9] = [[-40., 10.], [600., 30.], [5., -3.], [511.5, 511.5], [0., 0.], [250., 700.],
               [-200., -200.], [-0.5, 100.], [511.49, 3.]]
    basis = W.psf_grid_stamps()
    idx, w = W.bilinear_psf_weights(np.clip(pos[:, 0], 0, size - 1), np.clip(pos[:, 1], 0, size - 1),
                                    size, size)
    data = torch.zeros((size, size), dtype=torch.float32, device='cuda')
    fp = F.BatchForcedPhotometry(data, torch.ones_like(data), basis, idx, w, kind, pos, shape, tile=64)
    fp.plan()
    torch.cuda.synchronize()
    o = fp.order
    assert np.array_equal(np.sort(o), np.arange(n))
    P = F.plan_patches(np.asarray(kind)[o], pos[o], shape[o], (size, size), basis.shape[1:])
    assert np.array_equal(fp.win, P['win'])
    area = P['win'][:, 2] * P['win'][:, 3]
    sz = P['sz']
    want = np.where(sz == 0, 0, 1 + np.log2(np.maximum(sz // 32, 1)).astype(np.int64))
    want = np.where((sz > 0) & ((sz < 32) | (sz > 512)), 7, want)
    want = np.where(area == 0, 6, want)
    assert np.array_equal(fp.bucket, want)
    assert np.array_equal(fp.poff, np.cumsum(area) - area)
    border = fp.d_border.cpu().numpy()
    Jb = fp.d_jobs.cpu().numpy().view(FFT_JOB).reshape(-1)
    J = np.empty_like(Jb)
    J[border] = Jb
    ref = P['job']
    for f in ('sx', 'sy', 'mw', 'mh', 'K', 'mode', 'scale', 'amp'):
        assert np.array_equal(J[f], ref[f]), f
    assert np.array_equal(J['out_stride'], np.maximum(P['win'][:, 2], 1))
    assert np.array_equal(J['out_offset'], fp.poff)
    np.testing.assert_allclose(J['mux'], ref['mux'], rtol=0, atol=1e-12)
    np.testing.assert_allclose(J['muy'], ref['muy'], rtol=0, atol=1e-12)
    np.testing.assert_allclose(J['var'], ref['var'], rtol=2e-7, atol=0)
    # rank inside the bucket, in tile order
    bk = fp.bucket[border]
    starts = np.concatenate([[0], np.cumsum(np.bincount(bk, minlength=8))[:-1]])
    assert np.array_equal(Jb['psf'], np.arange(n) - starts[bk])
    assert np.all(np.diff(bk) >= 0) and np.all(np.diff(border)[np.diff(bk) == 0] > 0)


def _tiled_spd_system(rng, sizes, density=0.15):
    """A symmetric positive definite sparse matrix over tiles of the given
    sizes (couplings inside and between neighbouring tiles), as CSR."""
    n = int(sum(sizes))
    A = np.zeros((n, n))
    mask = rng.random((n, n)) < density
    mask = np.triu(mask | mask.T, 1)
    A[mask] = rng.normal(size=mask.sum())
    A = A + A.T
    A += np.diag(np.abs(A).sum(1) * 1.02 + rng.uniform(0.05, 2.0, n))   # strictly dominant
    rows, cols = np.nonzero(A)
    row_ptr = np.zeros(n + 1, np.int64)
    np.add.at(row_ptr, rows + 1, 1)
    row_ptr = np.cumsum(row_ptr)
    tile_start = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    block_off = np.concatenate([[0], np.cumsum(np.asarray(sizes, np.int64) ** 2)[:-1]])
    return A, row_ptr, cols.astype(np.int32), A[rows, cols], tile_start, block_off


def _device_pcg(A, row_ptr, col, vals, tile_start, block_off, b, tol, chunks):
    import ctypes
    import torch
    from tractor_amd import _lib
    lib = _lib.load()
    h = _lib.handle(0)
    dev = torch.device('cuda', 0)
    n, nt = len(b), len(tile_start) - 1
    max_n = int(np.diff(tile_start).max())
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    d_rp, d_col, d_vals, d_ts, d_bo, d_b = map(t, (row_ptr, col, vals, tile_start, block_off, b))
    blocks = torch.empty(int(block_off[-1] + (tile_start[-1] - tile_start[-2]) ** 2),
                         dtype=torch.float64, device=dev)
    info = torch.zeros(nt, dtype=torch.int32, device=dev)
    P = lambda x: ctypes.c_void_p(x.data_ptr())
    st = torch.cuda.current_stream(dev).cuda_stream
    assert lib.trh_fp_block_inv(h.value, P(d_rp), P(d_col), P(d_vals), P(d_ts), P(d_bo), nt,
                                 max_n, P(blocks), P(info), st) == 0
    x = torch.empty(n, dtype=torch.float64, device=dev)
    work = torch.empty(4 * n + 5 * nt + 2, dtype=torch.float64, device=dev)
    status = torch.empty(2, dtype=torch.int32, device=dev)
    it0 = 0
    for k in chunks:
        assert lib.trh_fp_pcg(h.value, n, P(d_rp), P(d_col), P(d_vals), P(d_ts), P(d_bo), nt,
                              max_n, P(blocks), P(d_b), tol, it0, k, P(x), P(work), P(status),
                              st) == 0, lib.trh_last_error()
        it0 += k
    torch.cuda.synchronize()
    assert int(info.abs().sum()) == 0
    return x.cpu().numpy(), status.cpu().numpy(), work[-2:].cpu().numpy()


@pytest.mark.parametrize('sizes', [[7, 12, 1, 30, 9], [64] * 6, [3]])
def test_device_pcg_matches_dense_solve(sizes):
    """trh_fp_pcg against numpy's dense FP64 solve; the stop test is the
    relative residual, so the error is bounded by tol * cond(A)."""
    rng = np.random.default_rng(len(sizes) * 7 + sizes[0])
    A, row_ptr, col, vals, ts, bo = _tiled_spd_system(rng, sizes)
    b = rng.normal(size=A.shape[0])
    ref = np.linalg.solve(A, b)
    print('pcg case', sizes, flush=True)
    # iterations enqueued in uneven chunks; launches after the stop are no-ops
    x, status, norms = _device_pcg(A, row_ptr, col, vals, ts, bo, b, 1e-12, [1, 5, 200])
    assert status[0] == 1 and 0 < status[1] <= 206
    cond = np.linalg.cond(A)
    assert np.max(np.abs(x - ref)) <= 1e-12 * cond * np.max(np.abs(ref)) * 10
    r = b - A @ x
    assert np.sqrt(norms[0]) <= 1e-12 * np.linalg.norm(b) * 1.0001
    assert norms[1] == pytest.approx(b @ b, rel=1e-12)
    # recursive residual stays close to the true one
    assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(b)
    # a single tile is solved exactly by its preconditioner in one step
    if len(sizes) == 1:
        assert status[1] == 1


def test_device_pcg_zero_rhs_and_iteration_cap():
    rng = np.random.default_rng(5)
    A, row_ptr, col, vals, ts, bo = _tiled_spd_system(rng, [20, 20, 20])
    x, status, norms = _device_pcg(A, row_ptr, col, vals, ts, bo, np.zeros(60), 1e-12, [4])
    assert status[0] == 1 and status[1] == 0 and np.all(x == 0) and norms[1] == 0
    b = rng.normal(size=60)
    x, status, _ = _device_pcg(A, row_ptr, col, vals, ts, bo, b, 0.0, [2])
    assert status[0] == 0 and status[1] == 2


def test_forced_photometry_psfex_matches_oracle():
    """Forced photometry with a PSFEx-style spatially varying PSF (degree-2
    polynomial over six eigen-PSFs; PixelizedPsfEx, psfex.py:294-346,
    481-520): the device's polynomial terms against PsfEx.polynomials, and
    the fluxes against the oracle rendering each source with psfex_at."""
    import torch
    import oracle as O
    from tractor_amd import forced as F, workloads as W
    size, n = 256, 120
    rng = np.random.default_rng(21)
    kind, pos, flux, shape = W.draw_forced_scene(n, size, rng=rng)
    grid = W.psf_grid_stamps().astype(np.float64)          # 9 stamps, 33 x 33
    mean = grid.mean(0)
    bases = np.stack([mean, 0.5 * (grid[5] - grid[3]), 0.1 * (grid[4] - mean),
                      0.5 * (grid[7] - grid[1]), 0.05 * (grid[8] - grid[0]),
                      0.1 * (grid[6] - mean)]).astype(np.float32)
    pe = dict(degree=2, x0=size / 2., y0=size / 2., xscale=size / 2., yscale=size / 2.)
    print('oracle patches', flush=True)
    patches = []
    for s in range(n):
        stamp = O.psfex_at(bases, pos[s, 0], pos[s, 1], 2, pe['x0'], pe['y0'], pe['xscale'],
                           pe['yscale'])
        patches.append(O.forced_unit_patch(int(kind[s]), pos[s, 0], pos[s, 1], shape[s],
                                           stamp.astype(np.float32), (size, size)))
    model = np.zeros((size, size))
    for pt, f in zip(patches, flux):
        if pt is not None:
            x0, y0, p = pt
            model[y0:y0 + p.shape[0], x0:x0 + p.shape[1]] += f * p
    data = (model + rng.normal(size=model.shape)).astype(np.float32)
    inverr = np.ones((size, size), np.float32)
    ref, iv_ref = O.forced_phot_lsq(patches, data.astype(np.float64), inverr.astype(np.float64))
    print('device forced photometry (PSFEx)', flush=True)
    fp = F.BatchForcedPhotometry(torch.from_numpy(data).cuda(), torch.from_numpy(inverr).cuda(),
                                 bases, None, None, kind, pos, shape, psfex=pe)
    got = fp.run(tol=1e-13)
    terms = np.empty((n, 6))
    terms[fp.order] = fp.d_pw.cpu().numpy()
    want = np.stack([O.psfex_terms(pos[s, 0], pos[s, 1], 2, pe['x0'], pe['y0'], pe['xscale'],
                                   pe['yscale']) for s in range(n)])
    np.testing.assert_allclose(terms, want, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(fp.iv, iv_ref, rtol=1e-4, atol=1e-12)
    sig = 1. / np.sqrt(np.maximum(iv_ref, 1e-300))
    err = np.abs(got - ref)
    ok = (err <= 1e-4 * np.abs(ref)) | (err <= 0.01 * sig)
    print('flux parity: %d/%d; max err/sigma %.3g' % (ok.sum(), n, np.max(err / sig)), flush=True)
    assert ok.all()


def test_device_catalog_gives_the_same_fit():
    """Sources and PSF weights passed as device tensors (source_catalog)
    give bit for bit the fluxes of the host-array path."""
    import torch
    from tractor_amd import forced as F, workloads as W
    size, n = 384, 600
    rng = np.random.default_rng(11)
    kind, pos, flux, shape = W.draw_forced_scene(n, size, rng=rng)
    basis = W.psf_grid_stamps()
    idx, w = W.bilinear_psf_weights(pos[:, 0], pos[:, 1], size, size)
    inverr = torch.ones((size, size), dtype=torch.float32, device='cuda')
    syn = F.BatchForcedPhotometry(torch.zeros_like(inverr), inverr, basis, idx, w, kind, pos,
                                  shape).plan().render()
    g = torch.Generator(device='cuda')
    g.manual_seed(3)
    data = syn.model_image(flux) + torch.randn((size, size), generator=g, device='cuda')
    a = F.BatchForcedPhotometry(data, inverr, basis, idx, w, kind, pos, shape)
    fa = a.run(tol=1e-12)
    cat = F.source_catalog(kind, pos, shape)
    b = F.BatchForcedPhotometry(data, inverr, basis, torch.from_numpy(idx).cuda(),
                                torch.from_numpy(w).cuda(), None, None, None, catalog=cat)
    fb = b.run(tol=1e-12)
    assert np.array_equal(np.asarray(fa), np.asarray(fb))
