"""Surrogate CNN initialiser for the batched fit.

A small CNN (four 3x3 / stride-2 convolutions with 32 channels, global pool,
MLP) reads a 32x32 cutout of each source's patch, normalised by its maximum,
and proposes the source type (point / exp / dev) and its starting parameters
(x, y, flux, log r_e, ee1, ee2).  Every layer runs in libtractor_hip: BF16
MFMA convolutions with FP32 accumulation, forward and backward, FP32 head and
Adam.  Training cutouts are rendered by the library's own renderer
(BatchFitter.render_models) from the batched-benchmark source distribution,
with Moffat PSFs of FWHM 3-6 px and unit-variance noise.

The fit that follows is the plain batched LM fit; `fit_with_fallback`
re-fits from the standard start every source whose surrogate-started fit
ends worse than the plain fit of the same source (final chi^2 above the
plain fit's by more than 1e-4 relative), or, without the plain fit's chi^2,
fails a chi-square self-check; the lower chi^2 is kept.
"""
import ctypes
import math
import time

import numpy as np
import torch

from . import _lib
from ._lib import check, ptr
from . import batch as Bt
from . import workloads as W

CIN = (1, 32, 32, 32)
C = 32
HEAD_PARAMS = 64 * 32 + 64 + 9 * 64 + 9
HEAD_SCR = 9 + 64 + 64 + 32
CUT = 32


def _layout():
    off_w, off_b = [], []
    o = 0
    for cin in CIN:
        off_w.append(o)
        o += 32 * 9 * cin
        off_b.append(o)
        o += 32
    return off_w, off_b, o, o + HEAD_PARAMS


OFF_W, OFF_B, OFF_HEAD, NPARAMS = _layout()


def init_params(seed=0):
    """He-normal conv and head weights, zero biases (FP32 flat vector)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAMS, np.float32)
    for l, cin in enumerate(CIN):
        n = 32 * 9 * cin
        p[OFF_W[l]:OFF_W[l] + n] = rng.normal(0, math.sqrt(2. / (9 * cin)), n)
    h = OFF_HEAD
    p[h:h + 2048] = rng.normal(0, math.sqrt(2. / 32), 2048)
    h2 = h + 2048 + 64
    p[h2:h2 + 576] = rng.normal(0, math.sqrt(1. / 64), 576)
    return p


def unpack_params(p):
    """Flat vector -> dict of the oracle's named arrays (cw<l>, cb<l>, hW1, ...)."""
    P = {}
    for l, cin in enumerate(CIN):
        P['cw%d' % l] = p[OFF_W[l]:OFF_W[l] + 32 * 9 * cin].reshape(32, 9, cin)
        P['cb%d' % l] = p[OFF_B[l]:OFF_B[l] + 32]
    h = OFF_HEAD
    P['hW1'] = p[h:h + 2048].reshape(64, 32)
    P['hb1'] = p[h + 2048:h + 2112]
    P['hW2'] = p[h + 2112:h + 2688].reshape(9, 64)
    P['hb2'] = p[h + 2688:h + 2697]
    return P


class SurrogateCNN(object):
    """Device parameters, Adam state and activation buffers for batches of up
    to `max_batch` cutouts."""

    def __init__(self, params=None, max_batch=4096, device=0, seed=0):
        self.lib = _lib.load()
        self.h = _lib.handle(device)
        self.dev = torch.device('cuda', device)
        p = init_params(seed) if params is None else np.asarray(params, np.float32)
        assert p.shape == (NPARAMS,)
        d = self.dev
        self.p = torch.from_numpy(p.copy()).to(d)
        self.g = torch.zeros_like(self.p)
        self.m = torch.zeros_like(self.p)
        self.v = torch.zeros_like(self.p)
        self.t = 0
        self.wf = [torch.zeros(32 * (16 if cin == 1 else 9 * cin), dtype=torch.bfloat16, device=d)
                   for cin in CIN]
        self.wt = [None] + [torch.zeros(32 * 9 * 32, dtype=torch.bfloat16, device=d)
                            for _ in CIN[1:]]
        self.norm_mean = np.zeros(6, np.float32)
        self.norm_std = np.ones(6, np.float32)
        self._alloc(max_batch)
        self.stream = torch.cuda.current_stream(d).cuda_stream

    def _alloc(self, B):
        d = self.dev
        bf = torch.bfloat16
        self.B = B
        self.y = [torch.zeros(B * s * s * C, dtype=bf, device=d) for s in (16, 8, 4, 2)]
        self.dz = [torch.zeros(B * s * s * C, dtype=bf, device=d) for s in (16, 8, 4, 2)]
        self.out = torch.zeros(B * 9, dtype=torch.float32, device=d)
        self.loss = torch.zeros(1, dtype=torch.float32, device=d)
        self.scr = torch.zeros(B * HEAD_SCR, dtype=torch.float32, device=d)

    def params_numpy(self):
        return self.p.cpu().numpy()

    # ---- device passes --------------------------------------------------
    def pack(self):
        for l, cin in enumerate(CIN):
            check(self.lib.trh_cnn_pack(self.h.value, ptr(self.p) + 4 * OFF_W[l], cin,
                                        ptr(self.wf[l]), ptr(self.wt[l]), self.stream), 'cnn_pack')

    def forward(self, x, B):
        """x: bf16 device tensor holding B cutouts (B, 32, 32, 1)."""
        if B > self.B or x.numel() < B * CUT * CUT or x.dtype != torch.bfloat16:
            raise ValueError('forward: batch of %d exceeds the %d allocated, or bad input' % (
                B, self.B))
        lib, h = self.lib, self.h.value
        src = ptr(x)
        H = CUT
        for l, cin in enumerate(CIN):
            check(lib.trh_cnn_conv_fwd(h, src, ptr(self.wf[l]), ptr(self.p) + 4 * OFF_B[l],
                                       ptr(self.y[l]), B, H, H, cin, self.stream), 'cnn_conv_fwd')
            src = ptr(self.y[l])
            H //= 2

    def head(self, B, targets=None, labels=None, rw=None, tw=1.0):
        train = targets is not None
        check(self.lib.trh_cnn_head(self.h.value, ptr(self.y[3]), ptr(self.p) + 4 * OFF_HEAD, B,
                                    ptr(targets) if train else None,
                                    ptr(labels) if train else None,
                                    ptr(rw) if train else None, float(tw), ptr(self.out),
                                    ptr(self.loss) if train else None,
                                    ptr(self.dz[3]) if train else None,
                                    ptr(self.scr) if train else None,
                                    (ptr(self.g) + 4 * OFF_HEAD) if train else None, self.stream),
              'cnn_head')

    def backward(self, x, B):
        lib, h = self.lib, self.h.value
        sizes = (32, 16, 8, 4)   # input side of layer l
        for l in range(3, -1, -1):
            H = sizes[l]
            xin = ptr(x) if l == 0 else ptr(self.y[l - 1])
            check(lib.trh_cnn_conv_bwd_weight(h, ptr(self.dz[l]), xin, ptr(self.g) + 4 * OFF_W[l],
                                              ptr(self.g) + 4 * OFF_B[l], B, H, H, CIN[l],
                                              self.stream), 'cnn_bwd_weight')
            if l > 0:
                check(lib.trh_cnn_conv_bwd_data(h, ptr(self.dz[l]), ptr(self.wt[l]),
                                                ptr(self.y[l - 1]), ptr(self.dz[l - 1]), B, H, H,
                                                self.stream), 'cnn_bwd_data')

    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.):
        self.t += 1
        lr_t = lr * math.sqrt(1 - b2 ** self.t) / (1 - b1 ** self.t)
        check(self.lib.trh_cnn_adam(self.h.value, ptr(self.p), ptr(self.g), ptr(self.m),
                                    ptr(self.v), NPARAMS, lr_t, b1, b2, eps, wd, self.stream),
              'cnn_adam')

    def train_step(self, x, targets, labels, B, lr, rw, tw=1.0):
        self.g.zero_()
        self.loss.zero_()
        self.pack()
        self.forward(x, B)
        self.head(B, targets, labels, rw, tw)
        self.backward(x, B)
        self.adam(lr)

    def predict(self, X, n, batch=None):
        """Raw network outputs (n, 9) for the first n cutouts of X."""
        batch = batch or self.B
        self.pack()
        out = torch.empty((n, 9), dtype=torch.float32, device=self.dev)
        for s in range(0, n, batch):
            B = min(batch, n - s)
            self.forward(X[s:s + B], B)
            self.head(B)
            out[s:s + B] = self.out[:B * 9].view(B, 9)
        return out

    # ---- training --------------------------------------------------------
    def fit(self, X, T, L, epochs=3, batch=None, lr=2e-3, tw=1.0, seed=0, log=None):
        """Train on cutouts X (n, 32, 32, 1) bf16, raw targets T (n, 6) and
        labels L (n,) int32 (device tensors).  Targets are standardised with
        the training-set mean/std (shape targets over galaxies only)."""
        batch = batch or self.B
        if batch > self.B:
            raise ValueError('fit: batch %d exceeds max_batch %d' % (batch, self.B))
        n = X.shape[0]
        Tn = T.cpu().numpy()
        Ln = L.cpu().numpy()
        gal = Ln > 0
        self.norm_mean = Tn.mean(0).astype(np.float32)
        self.norm_std = Tn.std(0).astype(np.float32)
        self.norm_mean[3:] = Tn[gal, 3:].mean(0)
        self.norm_std[3:] = Tn[gal, 3:].std(0)
        self.norm_std[self.norm_std == 0] = 1.
        Ts = ((T - torch.from_numpy(self.norm_mean).to(self.dev)) /
              torch.from_numpy(self.norm_std).to(self.dev)).contiguous()
        rw = torch.ones(6, dtype=torch.float32, device=self.dev)
        g = torch.Generator(device=self.dev)
        g.manual_seed(seed)
        steps = epochs * (n // batch)
        k = 0
        hist = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g, device=self.dev)
            for s in range(0, n - batch + 1, batch):
                idx = perm[s:s + batch]
                xb = X[idx].contiguous()
                tb = Ts[idx].contiguous()
                lb = L[idx].contiguous()
                lr_k = lr * 0.5 * (1 + math.cos(math.pi * k / max(steps, 1)))
                self.train_step(xb, tb, lb, batch, lr_k, rw, tw)
                k += 1
                if k == 1 or k % 50 == 0 or k == steps:
                    lv = float(self.loss.item())
                    hist.append((k, lv))
                    if log:
                        log('cnn step %d/%d loss %.4f' % (k, steps, lv))
        return hist

    def decode(self, out, maxv, P):
        """Network outputs -> (kind, params) in patch pixels."""
        o = out.cpu().numpy().astype(np.float64)
        maxv = np.asarray(maxv, np.float64)
        P = np.asarray(P)
        t = o[:, :6] * self.norm_std + self.norm_mean
        kind = np.argmax(o[:, 6:9], axis=1).astype(np.int32)
        params = np.zeros((len(o), 6))
        params[:, 0] = P // 2 + 0.5 + t[:, 0]
        params[:, 1] = P // 2 + 0.5 + t[:, 1]
        params[:, 2] = np.exp(t[:, 2]) * maxv
        params[:, 3:6] = t[:, 3:6]
        params[kind == 0, 3:6] = 0.
        return kind, params


def cutouts(data, pix_off, P, device=0):
    """32x32 bf16 cutouts (n, 32, 32, 1) centred on each patch's (P/2, P/2),
    divided by their maximum; returns (X, maxv)."""
    lib = _lib.load()
    h = _lib.handle(device)
    n = len(P)
    dev = torch.device('cuda', device)
    X = torch.empty((n, CUT, CUT, 1), dtype=torch.bfloat16, device=dev)
    maxv = torch.empty(n, dtype=torch.float32, device=dev)
    d_off = torch.from_numpy(np.ascontiguousarray(pix_off, np.int64)).to(dev)
    d_P = torch.from_numpy(np.ascontiguousarray(P, np.int32)).to(dev)
    check(lib.trh_cnn_cutouts(h.value, ptr(data), ptr(d_off), ptr(d_P), n, ptr(X), ptr(maxv),
                              torch.cuda.current_stream(dev).cuda_stream), 'cnn_cutouts')
    return X, maxv


def targets_of(kind, P, truth, maxv):
    """Raw regression targets (dx, dy, log(flux / max), log r_e, ee1, ee2)."""
    P = np.asarray(P)
    T = np.zeros((len(kind), 6), np.float32)
    T[:, 0] = truth[:, 0] - P // 2 - 0.5
    T[:, 1] = truth[:, 1] - P // 2 - 0.5
    T[:, 2] = np.log(truth[:, 2]) - np.log(np.asarray(maxv, np.float64))
    T[:, 3:6] = truth[:, 3:6]
    return T


def make_training_set(n, chunk=16384, seed=7, fwhm=(3., 6.), device=0, log=None):
    """n cutouts rendered by the library from the batched-benchmark source
    distribution (30/40/30 point/exp/dev, P 16..128, flux 50..1e5, r_e
    0.5..10 px), one Moffat PSF (beta 3.5, FWHM ~ U(fwhm)) per chunk, noise
    sigma 1.  Returns device tensors (X bf16, T float32, L int32)."""
    dev = torch.device('cuda', device)
    rng = np.random.default_rng(seed)
    X = torch.empty((n, CUT, CUT, 1), dtype=torch.bfloat16, device=dev)
    T = torch.empty((n, 6), dtype=torch.float32, device=dev)
    L = torch.empty(n, dtype=torch.int32, device=dev)
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        kind, P, truth = W.draw_sources(m, rng=rng)
        psf = W.moffat_stamp(63, float(rng.uniform(*fwhm)), 3.5)
        fitter, data, inverr, pix_off = Bt.make_problem(kind, P, truth, psf,
                                                        int(rng.integers(1 << 31)))
        Xc, maxv = cutouts(data, pix_off, P, device)
        X[s:s + m] = Xc
        T[s:s + m] = torch.from_numpy(targets_of(kind, P, truth, maxv.cpu().numpy())).to(dev)
        L[s:s + m] = torch.from_numpy(kind.astype(np.int32)).to(dev)
        fitter.close()
        del data, inverr, Xc
        if log:
            log('training set %d/%d' % (s + m, n))
    torch.cuda.synchronize()
    return X, T, L


def selfcheck_bad(chisq, P, nparams, nsig=5.):
    """Chi-square self-check: a converged fit of a correct model has chi^2 ~
    N_pix - N_par with standard deviation sqrt(2 N_pix)."""
    npix = np.asarray(P, np.float64) ** 2
    return chisq > (npix - nparams) + nsig * np.sqrt(2. * npix)


def fit_with_fallback(net, data, inverr, pix_off, P, psf, plain_start, plain_kind=None,
                      max_iter=100, dchisq=0., cdi=Bt.PIXEL_CDI, type_conf=0.0, timing=None,
                      profile=None, plain_chisq=None, rel_tol=1e-4, overlap=False):
    """Surrogate-initialised batched fit.

    1. cutouts -> CNN -> proposed type and start;
    2. one batched LM fit from the proposals; a source whose type probability
       is below `type_conf` (default 0: never) is fit under its second most
       likely type too, in the same batch, and keeps the lower chi^2;
    3. fallback: with `plain_chisq` (the final chi^2 of the plain fit of the
       same sources), every source whose chi^2 exceeds it by more than
       `rel_tol` relative (or is not a number) is re-fit from the standard
       start (and, if given, the standard type); without it, the sources
       failing the chi-square self-check are.  The lower chi^2 is kept.
       With `plain_chisq`, `overlap` and no second type hypotheses, the
       re-fits run in the same batch as chained restarts
       (BatchFitter.set_chain): each source's re-fit starts on the device as
       soon as its proposal's fit has stopped worse than the plain fit, so
       the re-fits overlap the rest of the fit instead of following it; the
       results are those of the separate re-fit (every source's fit is
       independent of the others in its batch).  That pays when the sources
       that need a re-fit stop early; a proposal that runs to max_iter and
       ends worse starts its re-fit only then, which stretches the batch to
       up to 2 max_iter iterations, so the separate re-fit is the default.
    Returns dict(kind, params, chisq, iterations, fallback mask, timings)."""
    t = {}
    n = len(P)
    P = np.asarray(P)
    pix_off = np.asarray(pix_off)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    X, maxv = cutouts(data, pix_off, P)
    out = net.predict(X, n)
    kind, start = net.decode(out, maxv.cpu().numpy(), P)
    o = out.cpu().numpy()[:, 6:9].astype(np.float64)
    pr = np.exp(o - o.max(1, keepdims=True))
    pr /= pr.sum(1, keepdims=True)
    order = np.argsort(-pr, axis=1)
    amb = np.flatnonzero(pr[np.arange(n), order[:, 0]] < type_conf)
    k2 = order[amb, 1].astype(np.int32)
    s2 = start[amb].copy()
    if len(amb):
        # a second galaxy hypothesis starts from the proposed shape (a point
        # source proposal has none: use a round 1-px galaxy)
        pt = kind[amb] == 0
        s2[pt, 3:6] = (0., 0., 0.)
        s2[k2 == 0, 3:6] = 0.
    kind_all = np.concatenate([kind, k2]).astype(np.int32)
    start_all = np.concatenate([start, s2])
    src = np.concatenate([np.arange(n), amb])
    chained = overlap and plain_chisq is not None and not len(amb)
    if chained:
        # sources n..2n-1: the standard start and type, each chained to its
        # proposal's fit
        pk = np.asarray(plain_kind if plain_kind is not None else kind).astype(np.int32)
        kind_all = np.concatenate([kind_all, pk]).astype(np.int32)
        start_all = np.concatenate([start_all, np.asarray(plain_start, np.float64)])
        src = np.concatenate([src, np.arange(n)])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    f = Bt.BatchFitter(kind_all, P[src], pix_off[src], data, inverr, psf, cdi=cdi)
    if chained:
        f.set_chain(np.arange(n), n + np.arange(n), plain_chisq, rel_tol)
    f.set_params(start_all)
    if profile is not None:
        f.profile(True)
    it = f.fit(max_iter=max_iter, dchisq=dchisq, graph=profile is None)
    if profile is not None:
        profile['slots'] = f.profile_read()
        f.profile(False)
    st = f.status()
    pall = f.get_params()
    cstate = f.chain_state() if chained else None
    f.close()
    chisq = st['chisq'][:n].copy()
    iters = st['iterations'][:n].copy()
    params = pall[:n].copy()
    fk = kind.copy()
    if len(amb):
        c2 = st['chisq'][n:]
        better = c2 < chisq[amb]
        sel = amb[better]
        chisq[sel] = c2[better]
        params[sel] = pall[n:][better]
        fk[sel] = k2[better]
        iters[amb] = np.maximum(iters[amb], st['iterations'][n:])
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    if chained:
        # the restarts that ran: keep the lower chi^2 of the pair
        bad = cstate == 1
        idx = np.flatnonzero(bad)
        c2 = st['chisq'][n:][idx]
        better = c2 < chisq[idx]
        sel = idx[better]
        chisq[sel] = c2[better]
        params[sel] = pall[n:][idx][better]
        fk[sel] = pk[idx][better]
        iters[idx] += st['iterations'][n:][idx]
    elif plain_chisq is not None:
        bad = ~(chisq <= np.asarray(plain_chisq, np.float64) * (1. + rel_tol))
    else:
        bad = selfcheck_bad(chisq, P, np.where(fk == 0, 3, 6))
    nb = int(bad.sum())
    if nb and not chained:
        idx = np.flatnonzero(bad)
        kb = (np.asarray(plain_kind)[idx] if plain_kind is not None else fk[idx]).astype(np.int32)
        fb = Bt.BatchFitter(kb, P[idx], pix_off[idx], data, inverr, psf, cdi=cdi)
        fb.set_params(np.asarray(plain_start)[idx])
        fb.fit(max_iter=max_iter, dchisq=dchisq)
        sb = fb.status()
        pb = fb.get_params()
        fb.close()
        better = sb['chisq'] < chisq[idx]
        sel = idx[better]
        chisq[sel] = sb['chisq'][better]
        params[sel] = pb[better]
        fk[sel] = kb[better]
        iters[idx] += sb['iterations']
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    t.update(cnn_s=t1 - t0, fit_s=t2 - t1, fallback_s=t3 - t2, total_s=t3 - t0, lm_iterations=it,
             second_hypotheses=int(len(amb)), fallback_sources=nb, fallback_overlapped=chained)
    if timing is not None:
        timing.update(t)
    return dict(kind=fk, params=params, chisq=chisq, iterations=iters, fallback=bad,
                proposed_kind=kind, timings=t)
