"""Host planning of batched forced photometry (CPU only): patch windows
against the oracle's extent rules, overlap pairs against brute force, the
CSR pattern, and the oracle's least-squares solve against closed forms."""
import numpy as np
import pytest

import oracle as O
from tractor_amd import forced as F, workloads as W


def _scene(n, size, seed):
    rng = np.random.default_rng(seed)
    kind, pos, flux, shape = W.draw_forced_scene(n, size, rng=rng)
    # push some sources onto and past the edges
    pos[:10] = rng.uniform(-20, size + 20, (10, 2))
    return kind, pos, flux, shape


def test_windows_match_oracle_extents():
    size = 160
    kind, pos, flux, shape = _scene(120, size, 1)
    stamp = W.psf_grid_stamps()[4]
    P = F.plan_patches(kind, pos, shape, (size, size), stamp.shape)
    for s in range(len(kind)):
        if kind[s] > 0 and np.exp(shape[s, 0]) > 4:
            continue  # keep the FP64 renders small; the rule is the same
        ref = O.forced_unit_patch(int(kind[s]), pos[s, 0], pos[s, 1], shape[s], stamp, (size, size))
        x0, y0, w, h = P['win'][s]
        if ref is None:
            assert w == 0 or h == 0, s
            continue
        rx0, ry0, p = ref
        assert (x0, y0, w, h) == (rx0, ry0, p.shape[1], p.shape[0]), (s, kind[s])


def test_overlap_pairs_brute_force():
    rng = np.random.default_rng(2)
    n = 300
    win = np.zeros((n, 4), np.int64)
    win[:, 0] = rng.integers(0, 500, n)
    win[:, 1] = rng.integers(0, 500, n)
    win[:, 2] = rng.integers(0, 60, n)
    win[:, 3] = rng.integers(0, 60, n)
    pairs = F.overlap_pairs(win, cell=37)
    got = set(map(tuple, pairs.tolist()))
    ref = set((s, s) for s in range(n))
    x0, y0, x1, y1 = win[:, 0], win[:, 1], win[:, 0] + win[:, 2], win[:, 1] + win[:, 3]
    for s in range(n):
        for t in range(s + 1, n):
            if (max(x0[s], x0[t]) < min(x1[s], x1[t]) and max(y0[s], y0[t]) < min(y1[s], y1[t])):
                ref.add((s, t))
    assert got == ref


def test_csr_of_pairs_symmetric():
    rng = np.random.default_rng(3)
    n = 40
    off = {(int(a), int(b)) for a, b in rng.integers(0, n, (80, 2)) if a < b}
    pairs = np.array(sorted({(s, s) for s in range(n)} | off), np.int64)
    row_ptr, col, pos = F.csr_of_pairs(pairs, n)
    vals = np.zeros(len(col))
    v = rng.normal(size=len(pairs))
    vals[pos[:, 0]] = v
    vals[pos[:, 1]] = v
    D = np.zeros((n, n))
    for r in range(n):
        for k in range(row_ptr[r], row_ptr[r + 1]):
            D[r, col[k]] = vals[k]
        assert np.all(np.diff(col[row_ptr[r]:row_ptr[r + 1]]) > 0)
    ref = np.zeros((n, n))
    for (s, t), x in zip(pairs, v):
        ref[s, t] = ref[t, s] = x
    assert np.array_equal(D, ref)


def test_oracle_lsq_exact_model_and_dense():
    # noise-free data = sum f u: the least-squares fluxes are exactly f
    rng = np.random.default_rng(4)
    H = Wd = 40
    patches = []
    for s in range(6):
        x0, y0 = rng.integers(0, 30, 2)
        patches.append((int(x0), int(y0), rng.uniform(0, 1, (10, 10))))
    patches.append(None)
    f = rng.uniform(1, 10, 7)
    f[-1] = 0.
    data = np.zeros((H, Wd))
    for (x0, y0, p), fs in zip(patches[:-1], f[:-1]):
        data[y0:y0 + 10, x0:x0 + 10] += fs * p
    inverr = rng.uniform(0.5, 2., (H, Wd))
    got, iv = O.forced_phot_lsq(patches, data, inverr)
    assert np.allclose(got, f, rtol=1e-10, atol=1e-10)
    # with noise: matches a dense least-squares solve
    noisy = data + rng.normal(size=data.shape)
    got, iv = O.forced_phot_lsq(patches, noisy, inverr)
    A = np.zeros((H * Wd, 6))
    for s, (x0, y0, p) in enumerate(patches[:-1]):
        m = np.zeros((H, Wd))
        m[y0:y0 + 10, x0:x0 + 10] = p
        A[:, s] = (m * inverr).ravel()
    ref = np.linalg.lstsq(A, (noisy * inverr).ravel(), rcond=None)[0]
    assert np.allclose(got[:6], ref, rtol=1e-9, atol=1e-9)
    assert np.allclose(iv[:6], (A ** 2).sum(0))


def _two_source_scene(rng, H=24, W=30):
    """Two overlapping positive patches and one isolated one (with a zero
    pixel), and their model at fluxes (3, -2, 5)."""
    u1 = rng.uniform(0.1, 1.0, (7, 9))
    u2 = rng.uniform(0.1, 1.0, (6, 8))
    u3 = rng.uniform(0.1, 1.0, (5, 5))
    u3[2, 2] = 0.
    patches = [(2, 3, u1), (6, 5, u2), (20, 15, u3), None]
    counts = np.array([3., -2., 5., 1.])
    model = np.zeros((H, W))
    for (x0, y0, u), c in zip(patches[:3], counts):
        model[y0:y0 + u.shape[0], x0:x0 + u.shape[1]] += c * u
    return patches, counts, model


def _fitstats_literal(patches, counts, mod, img, ie, scale, sky):
    """The steps of Optimizer._get_fitstats (optimize.py:233-376) one pixel
    at a time, for one image: srcmod accumulated from the unit model times
    counts and divided by the source counts, the sky taken out of the model
    before the per-source sums, nz the pixels with srcmod != 0 and ie > 0."""
    H, W = img.shape
    chi = (img - mod) * ie                  # mod includes the sky here
    modns = mod - sky
    out = []
    for (pt, c) in zip(patches, counts):
        st = dict(prochi2=0., pronpix=0., promasked=0., num=0., den2=0., den=0., proflux=0.,
                  npix=0, fracin=0., counted=0)
        if pt is None or c == 0:
            out.append(st)
            continue
        x0, y0, u = pt
        srcmod = np.zeros((H, W))
        for y in range(u.shape[0]):
            for x in range(u.shape[1]):
                srcmod[y0 + y, x0 + x] += u[y, x] * c
        srcmod /= c
        pix = [(y, x) for y in range(y0, y0 + u.shape[0]) for x in range(x0, x0 + u.shape[1])]
        nz = [(y, x) for (y, x) in pix if srcmod[y, x] != 0 and ie[y, x] > 0]
        if not nz:
            out.append(st)
            continue
        for (y, x) in nz:
            a = abs(srcmod[y, x])
            st['prochi2'] += a * chi[y, x] ** 2
            st['pronpix'] += a
            st['num'] += abs(modns[y, x] / c - srcmod[y, x]) * a
            st['den2'] += srcmod[y, x] ** 2
            st['den'] += a / abs(c)
            st['proflux'] += abs((modns[y, x] - srcmod[y, x] * c) / scale) * a
            st['npix'] += 1
        for (y, x) in pix:
            if ie[y, x] == 0:
                st['promasked'] += abs(srcmod[y, x])
            st['fracin'] += abs(srcmod[y, x])
        st['counted'] = 1
        out.append(st)
    return out


def test_oracle_fitstats_literal_steps():
    """forced_fitstats against the reference's steps transcribed one pixel
    at a time (sky, masked pixels, negative counts, a zero pixel, a source
    without a patch)."""
    rng = np.random.default_rng(3)
    patches, counts, model = _two_source_scene(rng)
    sky, scale = 0.7, 1.5
    data = model + sky + rng.normal(size=model.shape)
    ie = rng.uniform(0.5, 2.0, model.shape)
    ie[4, 4:7] = 0.
    ie[17, 21] = 0.
    fs = O.forced_fitstats(patches, counts, model, data, ie, scale=scale, sky=sky)
    lit = _fitstats_literal(patches, counts, model + sky, data, ie, scale, sky)
    for s, st in enumerate(lit):
        ff = st['num'] / st['den2'] if st['counted'] else 0.
        np.testing.assert_allclose(
            [fs['prochi2'][s], fs['pronpix'][s], fs['promasked'][s], fs['profracflux'][s],
             fs['proflux'][s], fs['npix'][s], fs['fracin'][s]],
            [st['prochi2'], st['pronpix'], st['promasked'], ff / max(1., st['den']),
             st['proflux'], st['npix'], st['fracin'] / max(1., st['counted'])],
            rtol=1e-12, atol=1e-12)
    chi = (data - model - sky) * ie
    assert fs['imchisq'] == pytest.approx(np.sum(chi ** 2), rel=1e-12)
    assert fs['imnpix'] == np.sum(ie > 0)


def test_oracle_fitstats_closed_forms():
    """Exact model: zero profile chi^2; an isolated source sees no flux
    from others; a constant offset k with inverr e gives prochi2 = (k e)^2
    pronpix; masking moves profile weight from pronpix to promasked; the
    other-source flux of source 1 is |counts_2| sum u1 u2 / scale."""
    rng = np.random.default_rng(4)
    patches, counts, model = _two_source_scene(rng)
    ie = np.full(model.shape, 2.)
    fs = O.forced_fitstats(patches, counts, model, model, ie)
    assert np.allclose(fs['prochi2'][:3], 0.) and fs['imchisq'] == 0.
    u3 = patches[2][2]
    assert fs['proflux'][2] == pytest.approx(0., abs=1e-12)
    assert fs['profracflux'][2] == pytest.approx(0., abs=1e-12)
    assert fs['npix'][2] == u3.size - 1
    assert fs['pronpix'][2] == pytest.approx(u3.sum()) and fs['fracin'][2] == pytest.approx(u3.sum())
    assert fs['iv'][2] == pytest.approx(np.sum((u3 * 2.) ** 2))
    k = 0.25
    fk = O.forced_fitstats(patches, counts, model, model + k, ie)
    np.testing.assert_allclose(fk['prochi2'][:3], (k * 2.) ** 2 * fk['pronpix'][:3], rtol=1e-12)
    # overlap of sources 1 and 2: u1 at (2, 3) 9x7, u2 at (6, 5) 8x6
    (x1, y1, u1), (x2, y2, u2) = patches[0], patches[1]
    ov = np.zeros(model.shape)
    ov[y2:y2 + u2.shape[0], x2:x2 + u2.shape[1]] = u2
    dot = np.sum(u1 * ov[y1:y1 + u1.shape[0], x1:x1 + u1.shape[1]])
    fs2 = O.forced_fitstats(patches, counts, model, model, ie, scale=2.)
    assert fs2['proflux'][0] == pytest.approx(abs(counts[1]) * dot / 2., rel=1e-12)
    # masking
    ie2 = ie.copy()
    ie2[y1:y1 + 3, x1:x1 + 4] = 0.
    fm = O.forced_fitstats(patches, counts, model, model, ie2)
    assert fm['pronpix'][0] + fm['promasked'][0] == pytest.approx(u1.sum(), rel=1e-12)
    assert fm['promasked'][0] == pytest.approx(u1[:3, :4].sum(), rel=1e-12)
    assert fm['npix'][0] == u1.size - 12


def test_oracle_fitstats_scaling_and_sign():
    """Scaling data, model and counts by k scales prochi2 by k^2 and proflux
    by k and leaves profracflux's numerator term alone; flipping the sign of
    every count, the model and the data changes nothing."""
    rng = np.random.default_rng(5)
    patches, counts, model = _two_source_scene(rng)
    data = model + rng.normal(size=model.shape)
    ie = rng.uniform(0.5, 2.0, model.shape)
    a = O.forced_fitstats(patches, counts, model, data, ie)
    k = 3.
    b = O.forced_fitstats(patches, k * counts, k * model, k * data, ie)
    np.testing.assert_allclose(b['prochi2'], k ** 2 * a['prochi2'], rtol=1e-12)
    np.testing.assert_allclose(b['proflux'], k * a['proflux'], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(b['ffnum'], a['ffnum'], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(b['ffden'], a['ffden'] / k, rtol=1e-12)
    c = O.forced_fitstats(patches, -counts, -model, -data, ie)
    for key in ('prochi2', 'pronpix', 'promasked', 'profracflux', 'proflux', 'fracin'):
        np.testing.assert_allclose(c[key], a[key], rtol=1e-12, atol=1e-12)


def test_oracle_psfex_terms_order_and_values():
    """PSFEx's term order (x powers fastest within each y power; psfex.py:
    304-318) written out for degrees 0..3, at a point with distinct dx, dy."""
    x0, y0, xs, ys = 100., 50., 40., 20.
    x, y = 120., 56.
    dx, dy = (x - x0) / xs, (y - y0) / ys      # 0.5, 0.3
    want = {0: [1.],
            1: [1., dx, dy],
            2: [1., dx, dx * dx, dy, dx * dy, dy * dy],
            3: [1., dx, dx ** 2, dx ** 3, dy, dx * dy, dx ** 2 * dy, dy ** 2, dx * dy ** 2,
                dy ** 3]}
    for deg, w in want.items():
        np.testing.assert_allclose(O.psfex_terms(x, y, deg, x0, y0, xs, ys), w, rtol=1e-15)


def test_oracle_psfex_at_reproduces_a_polynomial_field():
    """Bases chosen so that pixel (i, j) of sum_t terms_t basis_t is a known
    polynomial of the position: the model evaluated at several positions
    reproduces it exactly (each basis image a unit impulse per term)."""
    deg = 2
    nb = 6
    bases = np.zeros((nb, 3, 3))
    for t in range(nb):
        bases[t].flat[t] = 1.
    rng = np.random.default_rng(2)
    for _ in range(5):
        x, y = rng.uniform(0, 200, 2)
        img = O.psfex_at(bases, x, y, deg, 100., 100., 50., 50.)
        dx, dy = (x - 100.) / 50., (y - 100.) / 50.
        np.testing.assert_allclose(img.flat[:nb], [1., dx, dx * dx, dy, dx * dy, dy * dy],
                                   rtol=1e-14, atol=1e-15)
        assert np.all(img.flat[nb:] == 0.)
