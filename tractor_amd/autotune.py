"""Autotuner of the render schedule.

Renders are bucketed by FFT size (the reference's size rule fixes which
bucket a source falls in, tractor/psf.py:209-212); within a bucket the
launch shape -- wavefronts per block and the cap on G rows kept in LDS per
pass, which trades phase-A recomputation against blocks per CU -- is free.
`tune_fft` times candidate shapes on representative jobs of each bucket (the
patch sizes the workload actually has, in the line-search chi^2 mode that
dominates a fit) and installs the fastest with trh_set_fft_config.  Results
can be cached per device in a JSON file.
"""
import json
import os

import numpy as np
import torch

from . import _lib, render as R
from ._lib import FFT_JOB, TRH_CHI2, check, ptr

SIZES = (32, 64, 128, 256, 512)


def candidates(sz):
    """(threads, rows) shapes worth timing for FFT size sz (rows 0 = as
    many as LDS allows)."""
    th = [128, 256, 512, 1024] if sz <= 64 else [256, 512, 1024]
    rows = [0]
    if sz == 128:   # from 256 on the rows live in global memory, one pass
        rows += [sz // 2, sz - sz // 16]
    return [(t, r) for t in th for r in rows]


def bucket_of_P(P, psf_shape):
    """FFT size of a P x P modelMask render centred in its patch
    (galaxy.py:337-351)."""
    hs = R.mm_halfsize(P // 2 + 0.5, P // 2 + 0.5, (0, 0, P, P), psf_shape)
    return R.fft_size(hs)


def _jobs(sz, Ps, n, psf_shape, rng):
    J = np.zeros(n, FFT_JOB)
    P = rng.choice(Ps, n)
    off = np.zeros(n, np.int64)
    off[1:] = np.cumsum(P.astype(np.int64) ** 2)[:-1]
    J['out_offset'] = off
    J['out_stride'] = P
    J['mw'] = P
    J['mh'] = P
    cx = R.pad_center(sz, psf_shape[1])
    J['sx'] = P // 2 - cx
    J['sy'] = P // 2 - cx
    J['K'] = 9
    J['mode'] = TRH_CHI2
    J['slot'] = np.arange(n)
    J['scale'] = 100.
    J['mux'] = rng.uniform(-0.5, 0.5, n)
    J['muy'] = rng.uniform(-0.5, 0.5, n)
    for k in range(9):
        J['amp'][:, k] = 1. / 9
        v = 0.5 * (k + 1)
        J['var'][:, k, 0] = v
        J['var'][:, k, 1] = 0.1 * v
        J['var'][:, k, 2] = 0.8 * v
    return J, int(np.sum(P.astype(np.int64) ** 2))


def tune_fft(psf_stamp, patch_sizes, device=0, njobs=8192, reps=3, cache=None, seed=0):
    """Time every candidate shape per FFT bucket present in `patch_sizes`
    and install the fastest.  Returns {sz: (threads, rows, ns_per_render)}."""
    lib = _lib.load()
    h = _lib.handle(device)
    dev = torch.device('cuda', device)
    name = torch.cuda.get_device_name(device)
    if cache and os.path.exists(cache):
        with open(cache) as fh:
            saved = json.load(fh).get(name)
        if saved:
            for sz, (t, r, _) in saved.items():
                check(lib.trh_set_fft_config(h.value, int(sz), int(t), int(r)), 'set_fft_config')
            return {int(k): tuple(v) for k, v in saved.items()}
    stamp = np.ascontiguousarray(psf_stamp, np.float32)
    spectra = R.PsfSpectra(stamp, dev)
    rng = np.random.default_rng(seed)
    Ps = np.unique(np.asarray(patch_sizes))
    buckets = {}
    for P in Ps:
        buckets.setdefault(bucket_of_P(int(P), stamp.shape), []).append(int(P))
    out = {}
    for sz, plist in sorted(buckets.items()):
        if sz not in SIZES:
            continue
        J, npix = _jobs(sz, np.array(plist), njobs, stamp.shape, rng)
        dj = R.jobs_to_device(J, dev)
        data = torch.randn(npix, device=dev)
        inv = torch.ones_like(data)
        chi2 = torch.zeros(njobs, dtype=torch.float64, device=dev)
        spec, _, _ = spectra.get(sz)
        best = None
        for t, r in candidates(sz):
            check(lib.trh_set_fft_config(h.value, sz, t, r), 'set_fft_config')

            def run():
                check(lib.trh_render_fft_batch(h.value, sz, ptr(dj), njobs, ptr(spec), None,
                                               ptr(data), ptr(inv), ptr(chi2), None), 'render')
            run()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                run()
            e1.record()
            torch.cuda.synchronize()
            ns = e0.elapsed_time(e1) / reps * 1e6 / njobs
            if best is None or ns < best[2]:
                best = (t, r, ns)
        check(lib.trh_set_fft_config(h.value, sz, best[0], best[1]), 'set_fft_config')
        out[sz] = best
        del data, inv, chi2, dj
    if cache:
        allc = {}
        if os.path.exists(cache):
            with open(cache) as fh:
                allc = json.load(fh)
        allc[name] = {str(k): list(v) for k, v in out.items()}
        os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
        with open(cache, 'w') as fh:
            json.dump(allc, fh, indent=1)
    return out
