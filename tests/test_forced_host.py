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
