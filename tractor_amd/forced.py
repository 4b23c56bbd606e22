"""Batched forced photometry on one image (fluxes only).

The device counterpart of Tractor.optimize_forced_photometry
(tractor/engine.py:228-238 -> Optimizer.forced_photometry,
tractor/optimize.py:16-145, with LsqrOptimizer's unscaled least-squares step,
tractor/lsqr_optimizer.py:11-251) for thousands of overlapping sources with
a spatially varying pixelized PSF:

1. per-source PSFs by bilinear interpolation of a PSF grid, applied to the
   stamps (point sources) and, since the Fourier transform is linear, to the
   grid's spectra (galaxies) -- trh_psf_combine;
2. unit-flux patches with the reference's no-modelMask extents
   (HoggGalaxy._getUnitFluxPatchSize, galaxy.py:678-691 and the FFT path,
   galaxy.py:336-508; PixelizedPSF.getPointSourcePatch, psf.py:137-207),
   rendered in batches -- trh_render_fft_batch / trh_render_point_batch;
3. the sparsity pattern of the normal equations (trh_fp_overlaps) and their
   values over overlapping pairs (FP32 products, FP64 sums) --
   trh_fp_pair_dots_csr / trh_fp_rhs;
4. an FP64 solve: conjugate gradients preconditioned by the exact inverses
   of the diagonal blocks of at most 64 consecutive sources in tile order --
   trh_fp_block_inv, trh_fp_pcg.

The result is the least-squares flux vector LSQR converges to; `iv` are the
inverse variances sum (u inverr)^2 (Optimizer._get_iv, optimize.py:378-423).
"""
import math
import time

import numpy as np
import torch

from . import _lib, render as R
from ._lib import FFT_JOB, TRH_STORE, check, ptr
from .mixture_profiles import get_exp_mixture, get_dev_mixture
from .workloads import KIND_POINT, KIND_EXP, KIND_DEV


def _pad_center(sz, p):
    sz = np.asarray(sz)
    return np.where(sz >= p, (sz - p) // 2 + p // 2, p // 2 - (p - sz) // 2)


def plan_patches(kind, pos, shape, img_shape, psf_shape):
    """Unit-flux patch windows and render jobs without a modelMask.

    Returns dict(win (n, 4) [x0, y0, w, h] inside the image, sz (n,) FFT size
    (0 for point sources), job (n,) trh_fft_job records, var/amp filled for
    galaxies)."""
    kind = np.asarray(kind)
    n = len(kind)
    H, W = img_shape
    ph, pw = psf_shape
    px = np.asarray(pos[:, 0], np.float64)
    py = np.asarray(pos[:, 1], np.float64)
    win = np.zeros((n, 4), np.int64)
    szs = np.zeros(n, np.int64)
    J = np.zeros(n, FFT_JOB)
    J['scale'] = 1.
    J['mode'] = TRH_STORE
    # ---- point sources (psf.py:137-207, no modelMask): the Lanczos-shifted
    # stamp at (round(px) - pw//2, round(py) - ph//2), clipped to the image
    pt = kind == KIND_POINT
    ix = np.round(px[pt]).astype(np.int64)
    iy = np.round(py[pt]).astype(np.int64)
    x0 = ix - pw // 2
    y0 = iy - ph // 2
    X0, Y0 = np.maximum(x0, 0), np.maximum(y0, 0)
    X1, Y1 = np.minimum(x0 + pw, W), np.minimum(y0 + ph, H)
    win[pt] = np.stack([X0, Y0, np.maximum(X1 - X0, 0), np.maximum(Y1 - Y0, 0)], 1)
    J['sx'][pt] = x0 - X0
    J['sy'][pt] = y0 - Y0
    J['mux'][pt] = px[pt] - ix
    J['muy'][pt] = py[pt] - iy
    # ---- galaxies: halfsize (galaxy.py:678-691, NullWCS pixscale 1), extent
    # via get_overlapping_region (patch.pyx:20-66), FFT job (galaxy.py:336-508)
    gal = ~pt
    gi = np.flatnonzero(gal)
    nre = np.where(kind[gal] == KIND_EXP, 4., 8.)
    re = np.exp(shape[gal, 0])
    gx, gy = px[gal], py[gal]
    hs = np.maximum(1., nre * re)
    imsize = np.maximum.reduce([gx, W - 1 - gx, gy, H - 1 - gy]) + 1
    hs = np.minimum(hs, imsize)
    hs = np.maximum(hs, math.hypot(ph / 2., pw / 2.))
    hs = np.ceil(hs).astype(np.int64)
    xlo, xhi = np.floor(gx - hs).astype(np.int64), np.ceil(gx + hs + 1).astype(np.int64)
    ylo, yhi = np.floor(gy - hs).astype(np.int64), np.ceil(gy + hs + 1).astype(np.int64)
    gx0, gx1 = np.maximum(xlo, 0), np.minimum(xhi, W - 1) + 1
    gy0, gy1 = np.maximum(ylo, 0), np.minimum(yhi, H - 1) + 1
    sz = (2 ** np.ceil(np.log2(hs * 2.))).astype(np.int64)
    cx, cy = _pad_center(sz, pw), _pad_center(sz, ph)
    dx, dy = gx - cx, gy - cy
    ix0, iy0 = np.round(dx).astype(np.int64), np.round(dy).astype(np.int64)
    WX0, WY0 = np.maximum(ix0, gx0), np.maximum(iy0, gy0)
    WX1, WY1 = np.minimum(ix0 + sz, gx1), np.minimum(iy0 + sz, gy1)
    empty = (xlo > W - 1) | (xhi < 0) | (ylo > H - 1) | (yhi < 0)
    w_ = np.where(empty, 0, np.maximum(WX1 - WX0, 0))
    h_ = np.where(empty, 0, np.maximum(WY1 - WY0, 0))
    win[gal] = np.stack([np.where(empty, 0, WX0), np.where(empty, 0, WY0), w_, h_], 1)
    J['sx'][gal] = ix0 - WX0
    J['sy'][gal] = iy0 - WY0
    J['mux'][gal] = dx - ix0
    J['muy'][gal] = dy - iy0
    szs[gal] = sz
    # sheared MoG in pixels: var' = T var T^T with T = cd^-1 (3600 I) times
    # the EllipseESoft basis (ellipses.py:126-155, 232-255); the exp and dev
    # profiles are circular (var_k = v_k I), so var'_k = v_k T T^T.
    ee = np.hypot(shape[gal, 1], shape[gal, 2])
    e = 1. - np.exp(-ee)
    th = np.arctan2(shape[gal, 2], shape[gal, 1]) / 2.
    ct, st = np.cos(th), np.sin(th)
    ab = np.where(e >= 1., 1000., np.minimum(1000., (1. + e) / np.maximum(1. - e, 1e-300)))
    T = np.empty((len(gi), 2, 2))
    T[:, 0, 0], T[:, 0, 1] = re * ct / ab, re * st
    T[:, 1, 0], T[:, 1, 1] = -re * st / ab, re * ct
    M = np.einsum('nij,nkj->nik', T, T)
    for kk, mix in ((KIND_EXP, get_exp_mixture()), (KIND_DEV, get_dev_mixture())):
        sel = kind[gal] == kk
        rows = gi[sel]
        K = len(mix.amp)
        v = mix.var[:, 0, 0]
        J['K'][rows] = K
        J['amp'][rows, :K] = mix.amp
        J['var'][rows, :K, 0] = v[None, :] * M[sel, 0, 0, None]
        J['var'][rows, :K, 1] = v[None, :] * M[sel, 0, 1, None]
        J['var'][rows, :K, 2] = v[None, :] * M[sel, 1, 1, None]
    J['mw'] = win[:, 2]
    J['mh'] = win[:, 3]
    return dict(win=win, sz=szs, job=J)


def overlap_pairs(win, cell=48):
    """All pairs s <= t (diagonal included) whose windows intersect.  Each
    source is listed in every grid cell its window touches; a candidate pair
    is kept only in the cell holding the top-left corner of the two windows'
    intersection, so no pair is produced twice."""
    n = len(win)
    x0, y0, w, h = (win[:, k].astype(np.int64) for k in range(4))
    x1, y1 = x0 + w, y0 + h
    ok = np.flatnonzero((w > 0) & (h > 0))
    cx0, cy0 = x0[ok] // cell, y0[ok] // cell
    ncx = (x1[ok] - 1) // cell - cx0 + 1
    ncy = (y1[ok] - 1) // cell - cy0 + 1
    cnt = ncx * ncy
    rep = np.repeat(np.arange(len(ok)), cnt)
    loc = np.arange(len(rep)) - np.repeat(np.cumsum(cnt) - cnt, cnt)
    ccx = cx0[rep] + loc % ncx[rep]
    ccy = cy0[rep] + loc // ncx[rep]
    NCX = int(ccx.max()) + 2 if len(ccx) else 1
    cid = ccy * NCX + ccx
    order = np.argsort(cid, kind='stable')
    src = ok[rep[order]]
    cid = cid[order]
    # runs of equal cell id; within a run of length g, element k pairs with
    # the g-1-k elements after it
    start = np.r_[0, np.flatnonzero(np.diff(cid)) + 1]
    glen = np.diff(np.r_[start, len(cid)])
    gpos = np.arange(len(cid)) - np.repeat(start, glen)
    after = np.repeat(glen, glen) - 1 - gpos
    a = np.repeat(np.arange(len(cid)), after)
    b = a + 1 + (np.arange(len(a)) - np.repeat(np.cumsum(after) - after, after))
    sa, sb = src[a], src[b]
    s, t = np.minimum(sa, sb), np.maximum(sa, sb)
    ix0, iy0 = np.maximum(x0[s], x0[t]), np.maximum(y0[s], y0[t])
    keep = ((ix0 < np.minimum(x1[s], x1[t])) & (iy0 < np.minimum(y1[s], y1[t])) &
            ((iy0 // cell) * NCX + ix0 // cell == cid[a]))
    off = np.stack([s[keep], t[keep]], 1)
    off = off[np.lexsort((off[:, 1], off[:, 0]))] if len(off) else off.reshape(0, 2)
    diag = np.stack([np.arange(n), np.arange(n)], 1)
    return np.concatenate([diag, off]).astype(np.int64)


def csr_of_pairs(pairs, n):
    """Symmetric CSR pattern of the pairs; returns (row_ptr, col, pos) with
    pos[p] = (index of (s, t), index of (t, s))."""
    s, t = pairs[:, 0], pairs[:, 1]
    offd = s != t
    rows = np.concatenate([s, t[offd]])
    cols = np.concatenate([t, s[offd]])
    order = np.argsort(rows * np.int64(n) + cols)
    inv = np.empty_like(order)
    inv[order] = np.arange(len(order))
    npair = len(pairs)
    pos = np.zeros((npair, 2), np.int64)
    pos[:, 0] = inv[:npair]
    pos[:, 1] = inv[:npair]
    pos[offd, 1] = inv[npair:]
    row_ptr = np.zeros(n + 1, np.int64)
    row_ptr[1:] = np.cumsum(np.bincount(rows, minlength=n))
    return row_ptr, cols[order].astype(np.int32), pos


class BatchForcedPhotometry(object):
    """Forced photometry of n sources on one image (device tensors data,
    inverr of shape (H, W), float32).  psf_basis: (G, ph, pw) stamps of the
    PSF grid; psf_idx / psf_w: (n, 4) bilinear weights of each source.
    kind: 0 point, 1 exp, 2 dev; pos: (n, 2) pixels; shape: (n, 3)
    EllipseESoft (log r_e, ee1, ee2) of the galaxies.  Sources are processed
    in tile order (tile x tile pixel cells; by default the tile holds about
    40 sources), Morton order inside a tile, so the preconditioner blocks (at
    most max_block consecutive sources of one tile) are contiguous and
    spatially compact; results come back in the caller's order."""

    def __init__(self, data, inverr, psf_basis, psf_idx, psf_w, kind, pos, shape, tile=None,
                 max_block=64, device=0):
        self.lib = _lib.load()
        self.h = _lib.handle(device)
        self.dev = torch.device('cuda', device)
        assert data.dtype == torch.float32 and inverr.dtype == torch.float32
        assert data.shape == inverr.shape and data.dim() == 2
        self.data = data.contiguous()
        self.inverr = inverr.contiguous()
        self.H, self.W = data.shape
        self.basis = np.ascontiguousarray(psf_basis, np.float32)
        G, self.ph, self.pw = self.basis.shape
        assert self.ph % 2 == 1 and self.pw % 2 == 1
        n = len(kind)
        self.n = n
        if tile is None:
            # about 40 sources per tile (multiple of 16 pixels, 32..256)
            tile = int(np.clip(round(math.sqrt(40. * self.H * self.W / max(n, 1)) / 16) * 16,
                               32, 256))
        ntx = (self.W + tile - 1) // tile
        self.tile, self.ntx, self.nty = tile, ntx, (self.H + tile - 1) // tile
        tid = (np.clip(pos[:, 1], 0, self.H - 1) // tile).astype(np.int64) * ntx + \
            (np.clip(pos[:, 0], 0, self.W - 1) // tile).astype(np.int64)
        # inside a tile, Morton order over 8x8 sub-cells, so that a run of
        # consecutive sources (a preconditioner block) is spatially compact
        cx = (np.clip(pos[:, 0], 0, self.W - 1) % tile * 8 // tile).astype(np.int64)
        cy = (np.clip(pos[:, 1], 0, self.H - 1) % tile * 8 // tile).astype(np.int64)
        mort = np.zeros(n, np.int64)
        for bit in range(3):
            mort |= ((cx >> bit) & 1) << (2 * bit) | ((cy >> bit) & 1) << (2 * bit + 1)
        self.order = np.argsort(tid * 64 + mort, kind='stable')
        o = self.order
        self.stile = tid[o].astype(np.int32)
        self.kind = np.asarray(kind)[o]
        self.pos = np.asarray(pos, np.float64)[o]
        self.shape = np.asarray(shape, np.float64)[o]
        self.psf_idx = np.ascontiguousarray(np.asarray(psf_idx)[o], np.int32)
        self.psf_w = np.ascontiguousarray(np.asarray(psf_w)[o], np.float32)
        ts = np.bincount(tid[o], minlength=int(tid.max()) + 1)
        ts = ts[ts > 0]
        # preconditioner blocks: each tile's sources, split into near-equal
        # runs of at most max_block (each block is factored in LDS)
        parts = -(-ts // max_block)
        ts = np.concatenate([np.full(k, c // k) + (np.arange(k) < c % k) for c, k in zip(ts, parts)])
        self.tile_start = np.zeros(len(ts) + 1, np.int32)
        self.tile_start[1:] = np.cumsum(ts)
        self.stream = torch.cuda.current_stream(self.dev).cuda_stream
        self.timing = {}

    # ------------------------------------------------------------------
    def plan(self):
        t0 = time.perf_counter()
        P = plan_patches(self.kind, self.pos, self.shape, (self.H, self.W), (self.ph, self.pw))
        self.win = P['win']
        self.sz = P['sz']
        self.jobs = P['job']
        area = self.win[:, 2] * self.win[:, 3]
        self.poff = np.zeros(self.n, np.int64)
        self.poff[1:] = np.cumsum(area)[:-1]
        self.npix = int(area.sum())
        self.jobs['out_offset'] = self.poff
        self.jobs['out_stride'] = np.maximum(self.win[:, 2], 1)
        self.pattern()
        sizes = np.diff(self.tile_start).astype(np.int64)
        self.block_off = np.zeros(len(sizes), np.int64)
        self.block_off[1:] = np.cumsum(sizes ** 2)[:-1]
        self.max_tile = int(sizes.max())
        self.timing['plan_s'] = time.perf_counter() - t0
        return self

    def pattern(self):
        """The sparsity pattern of A on the device (trh_fp_overlaps): CSR
        row_ptr (kept on the host too), sorted columns and the diagonal's
        position in each row.  Two windows can only share a pixel when the
        cells of the sources' (clipped) positions are within `reach` cells,
        reach from the farthest window edge from its source position."""
        lib, h, st = self.lib, self.h.value, self.stream
        n, win = self.n, self.win
        px = np.clip(self.pos[:, 0], 0, self.W - 1)
        py = np.clip(self.pos[:, 1], 0, self.H - 1)
        ne = (win[:, 2] > 0) & (win[:, 3] > 0)
        ext = np.maximum.reduce([px - win[:, 0], win[:, 0] + win[:, 2] - px,
                                 py - win[:, 1], win[:, 1] + win[:, 3] - py])
        emax = float(ext[ne].max()) if ne.any() else 0.
        self.reach = int(math.floor(2. * emax / self.tile)) + 1
        ncell = self.ntx * self.nty
        cnt = np.bincount(self.stile, minlength=ncell)
        first = np.zeros(ncell, np.int32)
        first[1:] = np.cumsum(cnt)[:-1]
        self.d_win = self._t(win.astype(np.int32))
        self.d_stile = self._t(self.stile)
        self.d_tfirst = self._t(first)
        self.d_tend = self._t((first + cnt).astype(np.int32))
        counts = torch.zeros(n, dtype=torch.int32, device=self.dev)
        args = (h, n, ptr(self.d_win), ptr(self.d_stile), self.ntx, self.nty, ptr(self.d_tfirst),
                ptr(self.d_tend), self.reach)
        check(lib.trh_fp_overlaps(*args, ptr(counts), None, None, None, st), 'fp_overlaps')
        self.row_ptr = np.zeros(n + 1, np.int64)
        self.row_ptr[1:] = np.cumsum(counts.cpu().numpy())
        self.nnz = int(self.row_ptr[-1])
        self.d_row_ptr = self._t(self.row_ptr)
        self.d_col = torch.empty(self.nnz, dtype=torch.int32, device=self.dev)
        self.d_diag = torch.empty(n, dtype=torch.int64, device=self.dev)
        check(lib.trh_fp_overlaps(*args, None, ptr(self.d_row_ptr), ptr(self.d_col),
                                  ptr(self.d_diag), st), 'fp_overlaps')
        return self

    def _t(self, a, dt=None):
        a = np.ascontiguousarray(a)
        t = torch.from_numpy(a).to(self.dev)
        return t if dt is None else t.to(dt)

    def render(self):
        """Per-source PSFs and unit-flux patches on the device."""
        lib, h, st = self.lib, self.h.value, self.stream
        self.patches = torch.zeros(max(self.npix, 1), dtype=torch.float32, device=self.dev)
        basis = self._t(self.basis.reshape(len(self.basis), -1))
        d_idx = self._t(self.psf_idx)
        d_w = self._t(self.psf_w)
        # point sources: interpolated stamps, then one batched Lanczos render
        pt = np.flatnonzero((self.kind == KIND_POINT) & (self.win[:, 2] > 0) & (self.win[:, 3] > 0))
        if len(pt):
            stamps = torch.empty((len(pt), self.ph * self.pw), dtype=torch.float32, device=self.dev)
            ip = self._t(pt.astype(np.int64))
            for c0 in range(0, len(pt), 65535):
                c1 = min(c0 + 65535, len(pt))
                # keep the gathered arguments alive until the launch is queued
                wq = d_w[ip[c0:c1]].contiguous()
                iq = d_idx[ip[c0:c1]].contiguous()
                check(lib.trh_psf_combine(h, ptr(basis), len(self.basis), self.ph * self.pw,
                                          ptr(wq), ptr(iq),
                                          ptr(stamps) + 4 * c0 * self.ph * self.pw, c1 - c0, st),
                      'psf_combine')
            J = self.jobs[pt].copy()
            J['psf'] = np.arange(len(pt))
            dj = R.jobs_to_device(J, self.dev)
            check(lib.trh_render_point_batch(h, ptr(dj), len(pt), int(J['mw'].max()), ptr(stamps),
                                             self.ph, self.pw, ptr(self.patches), None, None, None,
                                             st), 'render_point_batch')
            self._keep = (stamps, dj)
        # galaxies: per FFT size, the grid's spectra, interpolated per source
        gal = (self.kind != KIND_POINT) & (self.win[:, 2] > 0) & (self.win[:, 3] > 0)
        keep = []
        for sz in np.unique(self.sz[gal]):
            sz = int(sz)
            gi = np.flatnonzero(gal & (self.sz == sz))
            nv = sz // 2 + 1
            spec_basis = torch.empty((len(self.basis), nv * sz * 2), dtype=torch.float32,
                                     device=self.dev)
            import ctypes
            cx, cy = ctypes.c_int(), ctypes.c_int()
            dst = self._t(self.basis)
            check(lib.trh_psf_spectrum_batch(h, ptr(dst), len(self.basis), self.ph, self.pw, sz,
                                             ptr(spec_basis), ctypes.byref(cx), ctypes.byref(cy),
                                             st), 'psf_spectrum_batch')
            spec = torch.empty((len(gi), nv * sz * 2), dtype=torch.float32, device=self.dev)
            ig = self._t(gi.astype(np.int64))
            for c0 in range(0, len(gi), 65535):
                c1 = min(c0 + 65535, len(gi))
                wq = d_w[ig[c0:c1]].contiguous()
                iq = d_idx[ig[c0:c1]].contiguous()
                check(lib.trh_psf_combine(h, ptr(spec_basis), len(self.basis), nv * sz * 2,
                                          ptr(wq), ptr(iq), ptr(spec) + 4 * c0 * nv * sz * 2,
                                          c1 - c0, st), 'psf_combine')
            J = self.jobs[gi].copy()
            J['psf'] = np.arange(len(gi))
            dj = R.jobs_to_device(J, self.dev)
            check(lib.trh_render_fft_batch(h, sz, ptr(dj), len(gi), ptr(spec), ptr(self.patches),
                                           None, None, None, st), 'render_fft_batch')
            keep.append((spec_basis, spec, dj, dst))
        self._keep_g = keep
        return self

    def assemble(self, mod0=None):
        """A (CSR values) and b on the device."""
        lib, h, st = self.lib, self.h.value, self.stream
        self.d_poff = self._t(self.poff)
        self.vals = torch.zeros(self.nnz, dtype=torch.float64, device=self.dev)
        self.b = torch.zeros(self.n, dtype=torch.float64, device=self.dev)
        check(lib.trh_fp_pair_dots_csr(h, self.n, ptr(self.d_row_ptr), ptr(self.d_col),
                                       ptr(self.d_win), ptr(self.d_poff), ptr(self.patches),
                                       ptr(self.inverr), self.W, ptr(self.vals), st),
              'fp_pair_dots_csr')
        check(lib.trh_fp_rhs(h, self.n, ptr(self.d_win), ptr(self.d_poff), ptr(self.patches),
                             ptr(self.data), ptr(mod0), ptr(self.inverr), self.W, ptr(self.b), st),
              'fp_rhs')
        return self

    def solve(self, tol=1e-12, maxiter=1000, chunk=8):
        """FP64 conjugate gradients preconditioned by the tile blocks' exact
        Cholesky factors (trh_fp_pcg: every scalar stays on the device).
        Iterations are enqueued `chunk` at a time; the stop flag of one chunk
        is read while the next one runs.  Returns fluxes in the caller's
        order."""
        lib, h, st = self.lib, self.h.value, self.stream
        d_rp = self.d_row_ptr
        d_col = self.d_col
        d_ts = self._t(self.tile_start)
        d_bo = self._t(self.block_off)
        nt = len(self.tile_start) - 1
        blocks = torch.empty(int(np.sum(np.diff(self.tile_start).astype(np.int64) ** 2)),
                             dtype=torch.float64, device=self.dev)
        info = torch.zeros(nt, dtype=torch.int32, device=self.dev)
        check(lib.trh_fp_block_inv(h, ptr(d_rp), ptr(d_col), ptr(self.vals), ptr(d_ts), ptr(d_bo),
                                    nt, self.max_tile, ptr(blocks), ptr(info), st), 'fp_block_inv')
        n = self.n
        x = torch.empty(n, dtype=torch.float64, device=self.dev)
        work = torch.empty(4 * n + 5 * nt + 2, dtype=torch.float64, device=self.dev)
        status = torch.empty(2, dtype=torch.int32, device=self.dev)
        pending = []
        it0 = 0
        while True:
            if it0 < maxiter:
                k = min(chunk, maxiter - it0)
                check(lib.trh_fp_pcg(h, n, ptr(d_rp), ptr(d_col), ptr(self.vals), ptr(d_ts),
                                     ptr(d_bo), nt, self.max_tile, ptr(blocks), ptr(self.b),
                                     float(tol), it0, k, ptr(x), ptr(work), ptr(status), st),
                      'fp_pcg')
                it0 += k
                hs = torch.empty(2, dtype=torch.int32, pin_memory=True)
                hs.copy_(status, non_blocking=True)
                ev = torch.cuda.Event()
                ev.record()
                pending.append((hs, ev))
            if len(pending) > 1 or it0 >= maxiter:
                hs, ev = pending.pop(0)
                ev.synchronize()
                if int(hs[0]) != 0 or not pending:
                    break
        torch.cuda.synchronize(self.dev)
        self.status = int(hs[0])
        self.iterations = int(hs[1])
        rr, bb = work[-2:].cpu().tolist()
        self.resid = float(np.sqrt(rr / bb)) if bb > 0 else 0.0
        self.info = info.cpu().numpy()
        flux = np.empty(n)
        flux[self.order] = x.cpu().numpy()
        diag = self.vals[self.d_diag].cpu().numpy()
        iv = np.empty(n)
        iv[self.order] = diag
        self.flux, self.iv = flux, iv
        return flux

    def model_image(self, flux, out=None):
        """sum_s flux_s u_s on the device (flux in the caller's order)."""
        if out is None:
            out = torch.zeros((self.H, self.W), dtype=torch.float32, device=self.dev)
        f = self._t(np.asarray(flux, np.float32)[self.order])
        dwin = self._t(self.win.astype(np.int32))
        dpoff = self._t(self.poff)
        check(self.lib.trh_fp_add_patches(self.h.value, self.n, ptr(dwin), ptr(dpoff),
                                          ptr(self.patches), ptr(f), self.W, ptr(out), self.stream),
              'fp_add_patches')
        return out

    def run(self, tol=1e-12, maxiter=1000):
        """plan -> render -> assemble -> solve; wall times per stage (host
        planning included) in self.timing and device times per stage from
        HIP events in self.device_ms."""
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        self.plan()
        t1 = time.perf_counter()
        ev[0].record()
        self.render()
        ev[1].record()
        self.assemble()
        ev[2].record()
        f = self.solve(tol=tol, maxiter=maxiter)
        ev[3].record()
        torch.cuda.synchronize()
        t4 = time.perf_counter()
        self.device_ms = dict(render=ev[0].elapsed_time(ev[1]), assemble=ev[1].elapsed_time(ev[2]),
                              solve=ev[2].elapsed_time(ev[3]))
        self.timing.update(plan_s=t1 - t0, device_s=(t4 - t1), total_s=t4 - t0)
        return f

    def algorithmic_bytes(self):
        """HBM bytes each stage must move: render (patch writes, per-source
        spectrum or stamp reads), assemble (u_s, u_t and inverr over every
        overlap, u/data/inverr over every patch) and one PCG iteration (CSR
        values and columns, gathered x)."""
        area = self.win[:, 2] * self.win[:, 3]
        gal = (self.kind != KIND_POINT) & (area > 0)
        spec = np.sum(8. * (self.sz[gal] // 2 + 1) * self.sz[gal])
        render = 4. * area.sum() + spec + 4. * self.ph * self.pw * np.sum((~gal) & (area > 0))
        col = self.d_col.cpu().numpy().astype(np.int64)
        rows = np.repeat(np.arange(self.n), np.diff(self.row_ptr))
        up = col >= rows
        s, t = rows[up], col[up]
        x0 = np.maximum(self.win[s, 0], self.win[t, 0])
        x1 = np.minimum(self.win[s, 0] + self.win[s, 2], self.win[t, 0] + self.win[t, 2])
        y0 = np.maximum(self.win[s, 1], self.win[t, 1])
        y1 = np.minimum(self.win[s, 1] + self.win[s, 3], self.win[t, 1] + self.win[t, 3])
        ov = np.maximum(x1 - x0, 0) * np.maximum(y1 - y0, 0)
        assemble = 12. * ov.sum() + 12. * area.sum() + 16. * self.nnz
        pcg = 12. * self.nnz + 8. * 10 * self.n
        return dict(render=render, assemble=assemble, pcg_iteration=pcg)
