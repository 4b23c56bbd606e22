"""Batched forced photometry on one image (fluxes only).

The device counterpart of Tractor.optimize_forced_photometry
(tractor/engine.py:228-238 -> Optimizer.forced_photometry,
tractor/optimize.py:16-145, with LsqrOptimizer's unscaled least-squares step,
tractor/lsqr_optimizer.py:11-251) for thousands of overlapping sources with
a spatially varying pixelized PSF:

1. per-source PSFs as weighted sums of basis images -- bilinear
   interpolation of a PSF grid, or PSFEx's polynomial over its eigen-PSFs
   (PixelizedPsfEx, psfex.py:294-346, 481-520; trh_psfex_terms) -- applied
   to the stamps (point sources) and, since the Fourier transform is linear,
   to the bases' spectra (galaxies) -- trh_psf_combine;
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


FFT_SIZES = (32, 64, 128, 256, 512)
B_POINT, B_EMPTY, B_BAD, NBUCKET = 0, 6, 7, 8   # trh_fp_plan's buckets; 1..5 FFT sizes


_MIX_DEV = {}


def _mixture_table_on(dev):
    """mixture_table() on the device, uploaded once per device."""
    t = _MIX_DEV.get(dev)
    if t is None:
        t = _MIX_DEV[dev] = torch.from_numpy(mixture_table()).to(dev)
    return t


def source_catalog(kind, pos, shape, device=0):
    """The (n, 6) float64 device catalogue BatchForcedPhotometry reads (x, y,
    kind, log r_e, ee1, ee2 per source): build it once and pass it as
    `catalog` to keep the sources resident across fits."""
    n = len(kind)
    src = np.empty((n, 6))
    src[:, 0:2] = pos
    src[:, 2] = kind
    src[:, 3:6] = shape
    return torch.from_numpy(src).to(torch.device('cuda', device))


def mixture_table():
    """The exp and dev profiles as trh_fp_plan's mix rows: K, amplitudes,
    variances (circular, in units of r_e^2)."""
    t = np.zeros((2, 1 + 2 * _lib.TRH_MAXK))
    for row, mix in enumerate((get_exp_mixture(), get_dev_mixture())):
        K = len(mix.amp)
        t[row, 0] = K
        t[row, 1:1 + K] = mix.amp
        t[row, 1 + _lib.TRH_MAXK:1 + _lib.TRH_MAXK + K] = mix.var[:, 0, 0]
    return t


class FitStats(object):
    """Fit statistics of a forced-photometry fit (the attributes of
    Optimizer._get_fitstats's FitStats, optimize.py:246-300)."""
    pass


class BatchForcedPhotometry(object):
    """Forced photometry of n sources on one image (device tensors data,
    inverr of shape (H, W), float32).  psf_basis: (G, ph, pw) stamps of the
    PSF grid; psf_idx / psf_w: (n, q) basis indices and weights of each
    source (bilinear: q = 4).  Or psfex = dict(degree, x0, y0, xscale,
    yscale): psf_basis holds PSFEx's eigen-PSFs and each source's weights are
    its polynomial terms at its position (psf_idx, psf_w unused).
    kind: 0 point, 1 exp, 2 dev; pos: (n, 2) pixels; shape: (n, 3)
    EllipseESoft (log r_e, ee1, ee2) of the galaxies.  Sources are processed
    in tile order (tile x tile pixel cells; by default the tile holds about
    40 sources), Morton order inside a tile, so the preconditioner blocks (at
    most max_block consecutive sources of one tile) are contiguous and
    spatially compact; results come back in the caller's order.  The whole
    plan (order, windows, render jobs, sparsity pattern, blocks) is built on
    the device; win, poff, order, row_ptr and tile_start are host copies made
    on first use."""

    def __init__(self, data, inverr, psf_basis, psf_idx, psf_w, kind, pos, shape, tile=None,
                 max_block=64, device=0, psfex=None, catalog=None):
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
        # the sources: host arrays, or a device catalogue (source_catalog);
        # psf_idx / psf_w may be device tensors too
        n = len(kind) if catalog is None else int(catalog.shape[0])
        assert n > 0
        self.n = n
        if tile is None:
            # about 40 sources per tile (multiple of 16 pixels, 32..256)
            tile = int(np.clip(round(math.sqrt(40. * self.H * self.W / n) / 16) * 16, 32, 256))
        self.tile = tile
        self.ntx, self.nty = (self.W + tile - 1) // tile, (self.H + tile - 1) // tile
        self.max_block = max_block
        if catalog is None:
            catalog = source_catalog(kind, pos, shape, device)
        assert catalog.dtype == torch.float64 and catalog.shape == (n, 6)
        self._d_src = catalog.to(self.dev).contiguous()
        self.psfex = psfex
        if psfex is None:
            def dev_t(a, dt, npdt):
                if isinstance(a, torch.Tensor):
                    return a.to(self.dev, dt).reshape(n, -1).contiguous()
                return torch.from_numpy(np.ascontiguousarray(a, npdt).reshape(n, -1)).to(self.dev)
            self._d_pidx0 = dev_t(psf_idx, torch.int32, np.int32)
            self._d_pw0 = dev_t(psf_w, torch.float32, np.float32)
            self.nq = self._d_pw0.shape[1]
        else:
            d = int(psfex['degree'])
            self.nq = (d + 1) * (d + 2) // 2
            assert self.nq == G, 'PSFEx degree %d needs %d bases, got %d' % (d, self.nq, G)
        self.stream = torch.cuda.current_stream(self.dev).cuda_stream
        self.timing = {}
        self._host = {}

    # ------------------------------------------------------------------
    def plan(self):
        """Sort keys -> tile order (stable sort) -> per-source windows, render
        buckets and jobs (trh_fp_plan) -> bucketed jobs, patch offsets,
        preconditioner blocks and the sparsity pattern.  Two small reads
        come back to the host: the sizes to allocate, then nnz."""
        t0 = time.perf_counter()
        lib, h, st, dev, n = self.lib, self.h.value, self.stream, self.dev, self.n
        i32, i64, f64 = torch.int32, torch.int64, torch.float64
        d_src = self._d_src
        key = torch.empty(n, dtype=i32, device=dev)
        check(lib.trh_fp_keys(h, n, ptr(d_src), self.H, self.W, self.tile, ptr(key), st),
              'fp_keys')
        order = torch.sort(key, stable=True).indices
        win = torch.empty((n, 4), dtype=i32, device=dev)
        bucket = torch.empty(n, dtype=i32, device=dev)
        stile = torch.empty(n, dtype=i32, device=dev)
        area = torch.empty(n, dtype=i64, device=dev)
        ext = torch.empty(n, dtype=f64, device=dev)
        jobs = torch.empty((n, FFT_JOB.itemsize), dtype=torch.uint8, device=dev)
        mix = _mixture_table_on(dev)
        check(lib.trh_fp_plan(h, n, ptr(order), ptr(d_src), ptr(key), self.H, self.W, self.ph,
                              self.pw, ptr(mix), ptr(win), ptr(bucket), ptr(stile), ptr(area),
                              ptr(ext), ptr(jobs), st), 'fp_plan')
        # cells, and the preconditioner blocks: each cell's run of sources
        # split into near-equal blocks of at most max_block
        ncell = self.ntx * self.nty
        ccount = torch.bincount(stile, minlength=ncell)
        nb = (ccount + self.max_block - 1) // self.max_block
        qd = ccount // nb.clamp(min=1)
        rd = ccount - qd * nb
        bcount = torch.bincount(bucket, minlength=NBUCKET)
        sizes = torch.cat([bcount.to(f64), area.sum().to(f64).view(1), ext.max().view(1),
                           nb.sum().to(f64).view(1), (qd + (rd > 0).to(i64)).max().to(f64).view(1),
                           (rd * (qd + 1) ** 2 + (nb - rd) * qd ** 2).sum().to(f64).view(1)])
        sizes = sizes.cpu().numpy()
        bc = sizes[:NBUCKET].astype(np.int64)
        self.npix, self.emax = int(sizes[NBUCKET]), float(sizes[NBUCKET + 1])
        self.ntile, self.max_tile, self.blocks_len = (int(v) for v in sizes[NBUCKET + 2:])
        self.bcount = bc
        self.bstart = np.concatenate([[0], np.cumsum(bc)[:-1]])
        # jobs grouped by bucket, each with its patch offset
        self.d_poff = torch.cumsum(area, 0) - area
        border = torch.sort(bucket, stable=True).indices
        d_bstart = torch.cumsum(bcount, 0) - bcount
        jobs_b = torch.empty_like(jobs)
        check(lib.trh_fp_jobs_bucketed(h, n, ptr(border), ptr(bucket), ptr(d_bstart),
                                       ptr(self.d_poff), ptr(jobs), ptr(jobs_b), st),
              'fp_jobs_bucketed')
        cell_of = torch.repeat_interleave(torch.arange(ncell, device=dev), nb,
                                          output_size=self.ntile)
        j = torch.arange(self.ntile, device=dev) - (torch.cumsum(nb, 0) - nb)[cell_of]
        bsize = qd[cell_of] + (j < rd[cell_of]).to(i64)
        self.d_ts = torch.cat([torch.zeros(1, dtype=i64, device=dev),
                               torch.cumsum(bsize, 0)]).to(i32)
        self.d_bo = torch.cumsum(bsize * bsize, 0) - bsize * bsize
        cfirst = torch.cumsum(ccount, 0) - ccount
        self.d_tfirst = cfirst.to(i32)
        self.d_tend = (cfirst + ccount).to(i32)
        self.d_order, self.d_win, self.d_stile, self.d_bucket = order, win, stile, bucket
        self.d_border, self.d_jobs = border, jobs_b
        if self.psfex is None:
            self.d_pidx = self._d_pidx0[order]
            self.d_pw = self._d_pw0[order]
        else:
            pe = self.psfex
            self.d_pw = torch.empty((n, self.nq), dtype=torch.float32, device=dev)
            self.d_pidx = torch.empty((n, self.nq), dtype=i32, device=dev)
            check(lib.trh_psfex_terms(h, n, ptr(order), ptr(d_src), int(pe['degree']),
                                      float(pe['x0']), float(pe['y0']), float(pe['xscale']),
                                      float(pe['yscale']), ptr(self.d_pw), ptr(self.d_pidx), st),
                  'psfex_terms')
        self._host = {}
        self.pattern()
        self.timing['plan_s'] = time.perf_counter() - t0
        return self

    def pattern(self):
        """The sparsity pattern of A on the device (trh_fp_overlaps): CSR
        row_ptr, sorted columns and the diagonal's position in each row.  Two
        windows can only share a pixel when the cells of the sources'
        (clipped) positions are within `reach` cells, reach from the farthest
        window edge from its source position."""
        lib, h, st, n = self.lib, self.h.value, self.stream, self.n
        self.reach = int(math.floor(2. * self.emax / self.tile)) + 1
        counts = torch.zeros(n, dtype=torch.int32, device=self.dev)
        args = (h, n, ptr(self.d_win), ptr(self.d_stile), self.ntx, self.nty, ptr(self.d_tfirst),
                ptr(self.d_tend), self.reach)
        check(lib.trh_fp_overlaps(*args, ptr(counts), None, None, None, None, st), 'fp_overlaps')
        self.d_row_ptr = torch.zeros(n + 1, dtype=torch.int64, device=self.dev)
        torch.cumsum(counts, 0, out=self.d_row_ptr[1:])
        self.nnz = int(self.d_row_ptr[-1])
        self.d_col = torch.empty(self.nnz, dtype=torch.int32, device=self.dev)
        self.d_diag = torch.empty(n, dtype=torch.int64, device=self.dev)
        self.d_erow = torch.empty(self.nnz, dtype=torch.int32, device=self.dev)
        check(lib.trh_fp_overlaps(*args, None, ptr(self.d_row_ptr), ptr(self.d_col),
                                  ptr(self.d_diag), ptr(self.d_erow), st), 'fp_overlaps')
        return self

    def _t(self, a, dt=None):
        a = np.ascontiguousarray(a)
        t = torch.from_numpy(a).to(self.dev)
        return t if dt is None else t.to(dt)

    def _hosted(self, name, fn):
        if name not in self._host:
            self._host[name] = fn()
        return self._host[name]

    # host copies of the device plan (tile order), for tests and byte counts
    win = property(lambda self: self._hosted('win', lambda: self.d_win.cpu().numpy()))
    poff = property(lambda self: self._hosted('poff', lambda: self.d_poff.cpu().numpy()))
    order = property(lambda self: self._hosted('order', lambda: self.d_order.cpu().numpy()))
    row_ptr = property(lambda self: self._hosted('row_ptr', lambda: self.d_row_ptr.cpu().numpy()))
    tile_start = property(lambda self: self._hosted('ts', lambda: self.d_ts.cpu().numpy()))
    bucket = property(lambda self: self._hosted('bucket', lambda: self.d_bucket.cpu().numpy()))

    def render(self):
        """Per-source PSFs and unit-flux patches on the device, one batched
        render per bucket of the bucketed jobs."""
        lib, h, st = self.lib, self.h.value, self.stream
        if self.bcount[B_BAD]:
            raise ValueError('%d galaxies need an FFT size outside 32..512' % self.bcount[B_BAD])
        self.patches = torch.zeros(max(self.npix, 1), dtype=torch.float32, device=self.dev)
        basis = self._t(self.basis.reshape(len(self.basis), -1))
        rec = FFT_JOB.itemsize
        keep = []

        def combine(src, length, b0, c, out):
            qi = self.d_border[b0:b0 + c]
            wq = self.d_pw[qi]
            iq = self.d_pidx[qi]
            keep.append((wq, iq))
            for c0 in range(0, c, 65535):
                c1 = min(c0 + 65535, c)
                check(lib.trh_psf_combine(h, ptr(src), len(self.basis), length,
                                          ptr(wq) + 4 * self.nq * c0, ptr(iq) + 4 * self.nq * c0,
                                          self.nq, ptr(out) + 4 * c0 * length, c1 - c0, st),
                      'psf_combine')
        # point sources: interpolated stamps, then one batched Lanczos render
        b0, c = int(self.bstart[B_POINT]), int(self.bcount[B_POINT])
        if c:
            stamps = torch.empty((c, self.ph * self.pw), dtype=torch.float32, device=self.dev)
            combine(basis, self.ph * self.pw, b0, c, stamps)
            check(lib.trh_render_point_batch(h, ptr(self.d_jobs) + rec * b0, c,
                                             ptr(stamps), self.ph, self.pw, ptr(self.patches),
                                             None, None, None, st), 'render_point_batch')
            keep.append(stamps)
        # galaxies: per FFT size, the grid's spectra, interpolated per source
        import ctypes
        for bi, sz in enumerate(FFT_SIZES, start=1):
            b0, c = int(self.bstart[bi]), int(self.bcount[bi])
            if not c:
                continue
            nv = sz // 2 + 1
            spec_basis = torch.empty((len(self.basis), nv * sz * 2), dtype=torch.float32,
                                     device=self.dev)
            cx, cy = ctypes.c_int(), ctypes.c_int()
            check(lib.trh_psf_spectrum_batch(h, ptr(basis), len(self.basis), self.ph, self.pw, sz,
                                             ptr(spec_basis), ctypes.byref(cx), ctypes.byref(cy),
                                             st), 'psf_spectrum_batch')
            spec = torch.empty((c, nv * sz * 2), dtype=torch.float32, device=self.dev)
            combine(spec_basis, nv * sz * 2, b0, c, spec)
            check(lib.trh_render_fft_batch(h, sz, ptr(self.d_jobs) + rec * b0, c, ptr(spec),
                                           ptr(self.patches), None, None, None, st),
                  'render_fft_batch')
            keep.append((spec_basis, spec))
        self._keep = (basis, keep)
        return self

    def assemble(self, mod0=None):
        """b and the weighted patches u inverr^2 (trh_fp_rhs), then A's CSR
        values from them (trh_fp_pair_dots_csr), on the device."""
        lib, h, st = self.lib, self.h.value, self.stream
        self.vals = torch.zeros(self.nnz, dtype=torch.float64, device=self.dev)
        self.b = torch.zeros(self.n, dtype=torch.float64, device=self.dev)
        self.wpatches = torch.empty_like(self.patches)
        check(lib.trh_fp_rhs(h, self.n, ptr(self.d_win), ptr(self.d_poff), ptr(self.patches),
                             ptr(self.data), ptr(mod0), ptr(self.inverr), self.W, ptr(self.b),
                             ptr(self.wpatches), st), 'fp_rhs')
        check(lib.trh_fp_pair_dots_csr(h, self.n, ptr(self.d_row_ptr), ptr(self.d_col),
                                       ptr(self.d_erow), ptr(self.d_win), ptr(self.d_poff),
                                       ptr(self.patches), ptr(self.wpatches), ptr(self.vals), st),
              'fp_pair_dots_csr')
        return self

    def solve(self, tol=1e-12, maxiter=1000, chunk=8):
        """FP64 conjugate gradients preconditioned by the exact inverses of
        the diagonal blocks (trh_fp_block_inv, trh_fp_pcg: every scalar stays
        on the device).  Iterations are enqueued `chunk` at a time; the stop
        flag of one chunk is read while the next one runs.  Returns fluxes in
        the caller's order."""
        lib, h, st = self.lib, self.h.value, self.stream
        d_rp, d_col, d_ts, d_bo = self.d_row_ptr, self.d_col, self.d_ts, self.d_bo
        nt = self.ntile
        blocks = torch.empty(max(self.blocks_len, 1), dtype=torch.float64, device=self.dev)
        info = torch.zeros(nt, dtype=torch.int32, device=self.dev)
        check(lib.trh_fp_block_inv(h, ptr(d_rp), ptr(d_col), ptr(self.vals), ptr(d_ts), ptr(d_bo),
                                   nt, self.max_tile, ptr(blocks), ptr(info), st), 'fp_block_inv')
        n = self.n
        x = torch.empty(n, dtype=torch.float64, device=self.dev)
        work = torch.empty(4 * n + 5 * nt + 2, dtype=torch.float64, device=self.dev)
        status = torch.empty(2, dtype=torch.int32, device=self.dev)
        pending = []
        it0 = 0
        res_prev = None   # relative residual at the end of the previous chunk read
        hold = False      # the chunk in flight is predicted to be the last one needed
        # per chunk: status and (r.r, b.b) copied to pinned host memory, and
        # an event; two slots (at most two chunks are in flight), allocated
        # once per solve
        hs_ring = torch.empty((2, 2), dtype=torch.int32, pin_memory=True)
        hn_ring = torch.empty((2, 2), dtype=torch.float64, pin_memory=True)
        ev_ring = [torch.cuda.Event(), torch.cuda.Event()]
        nchunk = 0
        while True:
            if it0 < maxiter and (not hold or not pending):
                k = min(chunk, maxiter - it0)
                check(lib.trh_fp_pcg(h, n, ptr(d_rp), ptr(d_col), ptr(self.vals), ptr(d_ts),
                                     ptr(d_bo), nt, self.max_tile, ptr(blocks), ptr(self.b),
                                     float(tol), it0, k, ptr(x), ptr(work), ptr(status), st),
                      'fp_pcg')
                it0 += k
                slot = nchunk & 1
                nchunk += 1
                hs, hn, ev = hs_ring[slot], hn_ring[slot], ev_ring[slot]
                hs.copy_(status, non_blocking=True)
                hn.copy_(work[-2:], non_blocking=True)
                ev.record()
                pending.append((hs, hn, ev))
            if len(pending) > 1 or it0 >= maxiter or hold:
                hs, hn, ev = pending.pop(0)
                ev.synchronize()
                if int(hs[0]) != 0 or (not pending and it0 >= maxiter):
                    break
                # the residual's decay over the chunk just read predicts the
                # iterations still needed; while the chunk in flight covers
                # them, no further chunk is enqueued (it would be `chunk`
                # pairs of launches that only return), and a wrong guess
                # costs one host round trip
                rr, bb = float(hn[0]), float(hn[1])
                res = np.sqrt(rr / bb) if bb > 0 else 0.0
                hold = False
                if pending and res_prev is not None and 0 < res < res_prev and res > tol:
                    rate = (res / res_prev) ** (1.0 / chunk)
                    hold = bool(np.log(tol / res) / np.log(rate) <= chunk - 0.5)
                res_prev = res
        torch.cuda.synchronize(self.dev)
        self.status = int(hs[0])
        self.iterations = int(hs[1])
        rr, bb = work[-2:].cpu().tolist()
        self.resid = float(np.sqrt(rr / bb)) if bb > 0 else 0.0
        self.info = info.cpu().numpy()
        out = torch.empty((2, n), dtype=torch.float64, device=self.dev)
        out[0, self.d_order] = x
        out[1, self.d_order] = self.vals[self.d_diag]
        self.flux, self.iv = out.cpu().numpy()
        return self.flux

    def model_image(self, flux, out=None):
        """sum_s flux_s u_s on the device (flux in the caller's order)."""
        if out is None:
            out = torch.zeros((self.H, self.W), dtype=torch.float32, device=self.dev)
        f = self._t(np.asarray(flux, np.float32))[self.d_order].contiguous()
        check(self.lib.trh_fp_add_patches(self.h.value, self.n, ptr(self.d_win), ptr(self.d_poff),
                                          ptr(self.patches), ptr(f), self.W, ptr(out), self.stream),
              'fp_add_patches')
        return out

    def fitstats(self, flux=None):
        """Optimizer._get_fitstats (optimize.py:233-376) of the fit on the
        device (trh_fp_fitstats, trh_fp_imstats): per source (caller order)
        prochi2, pronpix, promasked, profracflux, proflux, npix, fracin, and
        the image's imchisq, imnpix and sky (0: this class fits no sky).
        flux: counts per source in the caller's order (default: the solved
        fluxes)."""
        flux = self.flux if flux is None else np.asarray(flux, np.float64)
        lib, h, st, n = self.lib, self.h.value, self.stream, self.n
        model = self.model_image(flux)
        counts = self._t(np.asarray(flux, np.float64))[self.d_order].contiguous()
        out = torch.empty((n, 10), dtype=torch.float64, device=self.dev)
        im = torch.empty(2, dtype=torch.float64, device=self.dev)
        check(lib.trh_fp_fitstats(h, n, ptr(self.d_win), ptr(self.d_poff), ptr(self.patches),
                                  ptr(counts), ptr(model), ptr(self.data), ptr(self.inverr), self.W,
                                  1., 0., ptr(out), st), 'fp_fitstats')
        check(lib.trh_fp_imstats(h, self.H * self.W, ptr(model), ptr(self.data), ptr(self.inverr),
                                 0., ptr(im), st), 'fp_imstats')
        res = torch.empty_like(out)
        res[self.d_order] = out
        a = res.cpu().numpy()
        imv = im.cpu().numpy()
        fs = FitStats()
        fs.prochi2, fs.pronpix, fs.promasked = a[:, 0], a[:, 1], a[:, 2]
        fs.profracflux = a[:, 3] / np.maximum(1., a[:, 4])
        fs.proflux, fs.npix = a[:, 5], a[:, 6].astype(int)
        fs.fracin = a[:, 7] / np.maximum(1., a[:, 8])
        fs.imchisq = np.array([imv[0]])
        fs.imnpix = np.array([int(imv[1])])
        fs.sky = np.array([0.])
        return fs

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
        spectrum or stamp reads), assemble (u, data, inverr reads and the
        weighted patch's write over every patch, then w_s and u_t over every
        overlap, and the CSR values) and one PCG iteration (CSR values and
        columns, gathered x)."""
        area = self.win[:, 2] * self.win[:, 3]
        bk = self.bucket
        gal = (bk >= 1) & (bk <= len(FFT_SIZES))
        sz = 32 << np.where(gal, bk - 1, 0)
        spec = np.sum(8. * (sz[gal] // 2 + 1) * sz[gal])
        render = 4. * area.sum() + spec + 4. * self.ph * self.pw * np.sum(bk == B_POINT)
        col = self.d_col.cpu().numpy().astype(np.int64)
        rows = np.repeat(np.arange(self.n), np.diff(self.row_ptr))
        up = col >= rows
        s, t = rows[up], col[up]
        x0 = np.maximum(self.win[s, 0], self.win[t, 0])
        x1 = np.minimum(self.win[s, 0] + self.win[s, 2], self.win[t, 0] + self.win[t, 2])
        y0 = np.maximum(self.win[s, 1], self.win[t, 1])
        y1 = np.minimum(self.win[s, 1] + self.win[s, 3], self.win[t, 1] + self.win[t, 3])
        ov = np.maximum(x1 - x0, 0) * np.maximum(y1 - y0, 0)
        assemble = 8. * ov.sum() + 16. * area.sum() + 16. * self.nnz
        pcg = 12. * self.nnz + 8. * 10 * self.n
        return dict(render=render, assemble=assemble, pcg_iteration=pcg)
