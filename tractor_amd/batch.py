"""Batched single-source fitting on the GPU (trh_fitter_* in libtractor_hip).

The batched counterpart of running GpuOptimizer / ConstrainedOptimizer
(tractor/gpu_optimizer.py:184-811, tractor/constrained_optimizer.py:12-159)
on thousands of one-source Tractors: one persistent device state, one stream,
no host round trips inside an iteration.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check, ptr

NULLWCS_CDI = np.eye(2) * 3600.   # NullWCS(pixscale=1).cd_inverse (tractor/wcs.py:24-28)
PIXEL_CDI = np.eye(2)             # position parameters in pixels
# Active sources below which fit() replays the branched iteration graph
# (the sz = 128 renders on the fit's stream, the other buckets on a side
# stream; same results).  On MI355X, with the 16384-source batched-fit
# workload: 116.0 ms per fit never branching, 110.5 ms branching below
# 2048-8192 active sources, 110.7 ms branching throughout.
DEFAULT_BRANCH_BELOW = 4096


class BatchFitter(object):
    """S sources, each fit on its own P x P window of `data` / `inverr`
    (device float32 tensors): window s starts at pix_off[s] with row stride
    row_stride[s] (default P: patches back to back; an image's width: the
    sources' modelMasks in one shared image).

    kind: 0 PointSource, 1 ExpGalaxy, 2 DevGalaxy, 3 SersicGalaxy;  params
    (S, 6): x, y, flux, logre, ee1, ee2 (EllipseESoft) in patch pixels, and a
    seventh column, the Sersic index, when there are Sersic galaxies.
    cdi: 2x2 position-derivative matrix (tractor/ducks.py:413-423).
    """
    def __init__(self, kind, P, pix_off, data, inverr, psf_stamp, cdi=PIXEL_CDI, device=0,
                 row_stride=None):
        self.lib = _lib.load()
        self.h = _lib.handle(device)
        self.S = len(kind)
        self.kind = np.ascontiguousarray(kind, np.int32)
        self.P = np.ascontiguousarray(P, np.int32)
        self.pix_off = np.ascontiguousarray(pix_off, np.int64)
        if row_stride is None:
            row_stride = self.P
        self.row_stride = np.ascontiguousarray(np.broadcast_to(row_stride, self.P.shape), np.int32)
        if np.any(self.row_stride < self.P):
            raise ValueError('row_stride must be >= P')
        self.data = data
        self.inverr = inverr
        psf = np.ascontiguousarray(psf_stamp, np.float32)
        self.psf = psf
        cdi = np.ascontiguousarray(np.asarray(cdi, np.float64).ravel())
        assert data.dtype == torch.float32 and inverr.dtype == torch.float32
        assert data.is_cuda and inverr.is_cuda
        P64 = self.P.astype(np.int64)
        need = int(np.max(self.pix_off + (P64 - 1) * self.row_stride + P64))
        if data.numel() < need or inverr.numel() < need:
            raise ValueError('data/inverr hold %d/%d pixels, patches need %d' % (
                data.numel(), inverr.numel(), need))
        f = ctypes.c_void_p()
        check(self.lib.trh_fitter_create(self.h.value, ctypes.byref(f), self.S, ptr(self.kind),
                                         ptr(self.P), ptr(self.pix_off), ptr(self.row_stride),
                                         ptr(data), ptr(inverr),
                                         ptr(psf), psf.shape[0], psf.shape[1], ptr(cdi)),
              'trh_fitter_create')
        self.f = f
        self.n_chain = 0
        self.npar = 7 if np.any(self.kind == 3) else 6
        if self.npar == 7:
            from .sersic import fitter_table
            tab = fitter_table()
            check(self.lib.trh_fitter_set_sersic(f, ptr(tab), len(tab)), 'trh_fitter_set_sersic')
        self.stream = torch.cuda.current_stream().cuda_stream
        self.set_branching(DEFAULT_BRANCH_BELOW)

    def set_hybrid(self, amp=None, mean=None, var=None, inner_nsigma=3., outer_nsigma=4.):
        """Hybrid PSF (HybridPixelizedPSF, tractor/psf.py:537-625): galaxy
        components too wide for the FFT box go to pixel space, convolved with
        the PSF mixture amp[K], mean[K, 2], var[K, 2, 2] (pixels); see
        trh_fitter_set_hybrid.  No arguments: back to the pure pixelized PSF."""
        if amp is None:
            check(self.lib.trh_fitter_set_hybrid(self.f, 0, None, None, None, 3., 4.),
                  'trh_fitter_set_hybrid')
            return
        amp = np.ascontiguousarray(amp, np.float64).ravel()
        K = len(amp)
        mean = np.ascontiguousarray(np.asarray(mean, np.float64).reshape(K, 2))
        v = np.asarray(var, np.float64).reshape(K, 2, 2)
        v3 = np.ascontiguousarray(np.stack([v[:, 0, 0], v[:, 0, 1], v[:, 1, 1]], 1))
        check(self.lib.trh_fitter_set_hybrid(self.f, K, ptr(amp), ptr(mean), ptr(v3),
                                             float(inner_nsigma), float(outer_nsigma)),
              'trh_fitter_set_hybrid')

    def set_branching(self, below_active):
        """Replay the branched iteration graph (render buckets of each phase
        on parallel streams) once fewer than `below_active` sources are left
        (trh_fitter_set_branching; 0: never).  Same results."""
        check(self.lib.trh_fitter_set_branching(self.f, int(below_active)), 'set_branching')

    def set_chain(self, first, second, ref_chisq, rel_tol=1e-4):
        """Chained restarts (trh_fitter_set_chain): source second[i] starts,
        from the parameters set for it, once source first[i] has stopped with
        chi^2 above ref_chisq[i] (1 + rel_tol); until then, and if it is not
        needed, it stays stopped.  Takes effect from the next set_params."""
        first = np.ascontiguousarray(first, np.int32)
        second = np.ascontiguousarray(second, np.int32)
        ref = np.ascontiguousarray(ref_chisq, np.float64)
        if not (len(first) == len(second) == len(ref)):
            raise ValueError('set_chain: first, second and ref_chisq differ in length')
        check(self.lib.trh_fitter_set_chain(self.f, len(first), ptr(first), ptr(second), ptr(ref),
                                            float(rel_tol), self.stream), 'set_chain')
        self.n_chain = len(first)

    def chain_state(self):
        """Per pair: 0 first source running, 1 second started, 2 not needed."""
        st = np.zeros(max(self.n_chain, 1), np.int32)
        check(self.lib.trh_fitter_chain_state(self.f, ptr(st), self.stream), 'chain_state')
        return st[:self.n_chain]

    def close(self):
        if self.f:
            self.lib.trh_fitter_destroy(self.f)
            self.f = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params):
        p = np.asarray(params, np.float64)
        if p.shape != (self.S, self.npar):
            raise ValueError('params: shape %s, want (%d, %d)' % (p.shape, self.S, self.npar))
        p7 = np.zeros((self.S, 7))
        p7[:, :self.npar] = p
        check(self.lib.trh_fitter_set_params(self.f, ptr(p7), self.stream), 'set_params')

    def get_params(self):
        p = np.zeros((self.S, 7))
        check(self.lib.trh_fitter_get_params(self.f, ptr(p), self.stream), 'get_params')
        return p[:, :self.npar].copy()

    def render_models(self, out, accumulate=False):
        """Models at the current parameters into `out` (laid out like data):
        stored, or added (overlapping windows sum)."""
        check(self.lib.trh_fitter_render_models(self.f, ptr(out), 1 if accumulate else 0,
                                                self.stream), 'render_models')

    def iterate(self, lam=0., dchisq=0., max_iter=100):
        check(self.lib.trh_fitter_iterate(self.f, float(lam), float(dchisq), int(max_iter),
                                          self.stream), 'iterate')

    def iterate_graph(self, n, lam=0., dchisq=0., max_iter=100):
        """n iterations replayed from a captured hipGraph of one iteration."""
        check(self.lib.trh_fitter_iterate_graph(self.f, float(lam), float(dchisq),
                                                int(max_iter), int(n), self.stream),
              'iterate_graph')

    def flags(self):
        """Per-source TRH_FIT_F_* bits (see include/tractor_hip.h)."""
        fl = np.zeros(self.S, np.int32)
        check(self.lib.trh_fitter_flags(self.f, ptr(fl), self.stream), 'flags')
        return fl

    def status(self):
        active = np.zeros(self.S, np.int32)
        iters = np.zeros(self.S, np.int32)
        chisq = np.zeros(self.S)
        dlnp = np.zeros(self.S)
        n = ctypes.c_int32()
        check(self.lib.trh_fitter_status(self.f, ptr(active), ptr(iters), ptr(chisq), ptr(dlnp),
                                         ctypes.byref(n), self.stream), 'status')
        return dict(active=active, iterations=iters, chisq=chisq, dlnp=dlnp, n_active=n.value)

    def n_active(self):
        n = ctypes.c_int32()
        check(self.lib.trh_fitter_status(self.f, None, None, None, None, ctypes.byref(n),
                                         self.stream), 'status')
        return n.value

    def timing(self):
        ms = np.zeros(6, np.float32)
        check(self.lib.trh_fitter_timing(self.f, ptr(ms)), 'timing')
        return ms

    def profile(self, enable=True):
        check(self.lib.trh_fitter_profile(self.f, 1 if enable else 0), 'profile')

    def profile_read(self):
        out = np.zeros((15, 4))
        check(self.lib.trh_fitter_profile_read(self.f, ptr(out)), 'profile_read')
        return out

    def fit(self, max_iter=100, dchisq=0., lam=0., check_every=None, graph=True):
        """Iterate until every source meets the optimize_loop stopping rule
        (sources that have stopped are skipped by every kernel, so extra
        iterations between checks change nothing).  graph=True replays a
        captured hipGraph of one iteration, check_every (default 4) at a
        time; graph=False launches eagerly (default check every iteration)."""
        if check_every is None:
            check_every = 4 if graph else 1
        if max_iter < 1:
            raise ValueError('max_iter must be >= 1')
        # a chained second source can start as late as its first's last
        # iteration and then runs max_iter of its own
        total = 2 * max_iter if self.n_chain else max_iter
        it = 0
        pending = False
        while it < total:
            k = min(check_every, total - it)
            if not graph:
                for _ in range(k):
                    self.iterate(lam, dchisq, max_iter)
                it += k
                if self.n_active() == 0:
                    break
                continue
            self.iterate_graph(k, lam, dchisq, max_iter)
            it += k
            # the check lags one chunk: the count of the previous chunk is
            # read while this one runs (trh_fitter_count_active), so the
            # device never waits for the host; a fit that has finished runs
            # one more chunk of iterations with nothing left to do
            if pending:
                n = self.count_read(wait=True)
                pending = False
                if n == 0:
                    break
            self.count_active()
            pending = True
        if pending:
            self.count_read(wait=True)
        return it

    def count_active(self):
        """Enqueue a count of the active sources (trh_fitter_count_active)."""
        check(self.lib.trh_fitter_count_active(self.f, self.stream), 'count_active')

    def count_read(self, wait=True):
        """The enqueued count (-1 without `wait` if it has not arrived)."""
        n = ctypes.c_int32()
        check(self.lib.trh_fitter_count_read(self.f, 1 if wait else 0, ctypes.byref(n)),
              'count_read')
        return n.value

    # ---- debugging views of device state (tests) ----
    def debug_arrays(self):
        r = ctypes.c_void_p()
        ro = ctypes.c_void_p()
        ne = ctypes.c_void_p()
        x = ctypes.c_void_p()
        check(self.lib.trh_fitter_debug(self.f, ctypes.byref(r), ctypes.byref(ro), ctypes.byref(ne),
                                        ctypes.byref(x)), 'debug')
        nr = np.select([self.kind == 3, self.kind > 0], [5, 4], 1) * self.P.astype(np.int64) ** 2
        total = int(nr.sum())
        renders = np.zeros(total, np.float32)
        neq = np.zeros((self.S, 40))
        X = np.zeros((self.S, 8))
        roff = np.zeros(self.S, np.int64)
        for dst, src, nb in [(renders, r, renders.nbytes), (neq, ne, neq.nbytes), (X, x, X.nbytes),
                             (roff, ro, roff.nbytes)]:
            check(self.lib.trh_copy_to_host(ptr(dst), src.value, nb, self.stream), 'copy')
        return dict(renders=renders, render_off=roff, neq=neq, X=X)


def make_problem(kind, P, truth, psf_stamp, noise_seed, device='cuda', sigma=1.0, cdi=PIXEL_CDI):
    """Synthesize data = render(truth) + N(0, sigma^2) on the device, inverr =
    1/sigma, and return (fitter, data, inverr, pix_off)."""
    from .workloads import patch_offsets
    pix_off, total = patch_offsets(np.asarray(P))
    data = torch.zeros(total, dtype=torch.float32, device=device)
    inverr = torch.full((total,), 1.0 / sigma, dtype=torch.float32, device=device)
    fitter = BatchFitter(kind, P, pix_off, data, inverr, psf_stamp, cdi=cdi)
    fitter.set_params(truth)
    fitter.render_models(data)
    g = torch.Generator(device=device)
    g.manual_seed(int(noise_seed))
    data.add_(torch.randn(total, generator=g, device=device, dtype=torch.float32) * sigma)
    torch.cuda.synchronize()
    return fitter, data, inverr, pix_off
