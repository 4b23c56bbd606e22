"""Host-side render planning: the reference's extent / FFT-size / shift rules,
turned into job descriptors for libtractor_hip.

The arithmetic here is index and shape bookkeeping only (what the reference
does in Python around its C calls); every pixel is computed on the GPU.
"""
import math

import numpy as np
import torch

from . import _lib
from ._lib import FFT_JOB, MOG_JOB, TRH_STORE, TRH_ADD, TRH_CHI2, TRH_MAXK, check, ptr

SUPPORTED_FFT_SIZES = (32, 64, 128, 256, 512)


def fft_size(radius):
    """PixelizedPSF.getFourierTransformSize (tractor/psf.py:209-212)."""
    return 2 ** int(np.ceil(np.log2(radius * 2.)))


def pad_center(sz, p):
    """Centre of a p-pixel stamp padded/cropped to sz (tractor/psf.py:225-238)."""
    if sz >= p:
        return (sz - p) // 2 + p // 2
    return p // 2 - (p - sz) // 2


def get_overlapping_region(xlo, xhi, xmin, xmax):
    """Inclusive-range overlap (tractor/patch.pyx:20-66)."""
    if xlo > xmax or xhi < xmin or xlo > xhi or xmin > xmax:
        return ([], [])
    xloclamp = max(xlo, xmin)
    Xlo = xloclamp - xlo
    xhiclamp = min(xhi, xmax)
    Xhi = Xlo + (xhiclamp - xloclamp)
    return (slice(xloclamp, xhiclamp + 1), slice(Xlo, Xhi + 1))


def mm_halfsize(px, py, mm, psf_shape):
    """FFT halfsize with a modelMask (tractor/galaxy.py:337-351)."""
    x0, y0, mw, mh = mm
    x1, y1 = x0 + mw, y0 + mh
    halfsize = max(mh / 2., mw / 2.)
    halfsize = max(halfsize, max(max(1 + px - x0, 1 + x1 - px),
                                 max(1 + py - y0, 1 + y1 - py)))
    psfh, psfw = psf_shape
    halfsize = max(halfsize, max(psfw / 2., psfh / 2.))
    return int(np.ceil(halfsize))


def galaxy_fft_job(amp, var_pix, px, py, psf_shape, mm=None, nomask_extent=None,
                   halfsize=None):
    """Plan one FFT-path render (tractor/galaxy.py:336-508).

    amp: (K,), var_pix: (K,2,2) sheared profile in pixels; mm: modelMask
    (x0, y0, w, h) or None with nomask_extent (x0, x1, y0, y1, halfsize);
    halfsize: the FFT halfsize to use with a modelMask instead of its rule
    (force_halfsize, galaxy.py:348-349).
    Returns (sz, job record, (X0, Y0, w, h)) where the last tuple is the
    output window in image coordinates.
    """
    K = len(amp)
    if K > TRH_MAXK:
        raise ValueError('at most %d MoG components per FFT render' % TRH_MAXK)
    psfh, psfw = psf_shape
    if mm is not None:
        x0, y0, mw, mh = mm
        halfsize = mm_halfsize(px, py, mm, psf_shape) if halfsize is None else halfsize
    else:
        x0, x1, y0, y1, halfsize = nomask_extent
    sz = fft_size(halfsize)
    cx, cy = pad_center(sz, psfw), pad_center(sz, psfh)
    dx, dy = px - cx, py - cy
    J = np.zeros((), FFT_JOB)
    if mm is not None:
        mux, muy = dx - x0, dy - y0
        sx, sy = int(np.round(mux)), int(np.round(muy))
        mux -= sx
        muy -= sy
        win = (x0, y0, mw, mh)
    else:
        ix0, iy0 = int(np.round(dx)), int(np.round(dy))
        mux, muy = dx - ix0, dy - iy0
        # clip G (placed at ix0, iy0) to [x0,x1) x [y0,y1) (galaxy.py:495-508)
        X0, Y0 = max(ix0, x0), max(iy0, y0)
        X1, Y1 = min(ix0 + sz, x1), min(iy0 + sz, y1)
        sx, sy = ix0 - X0, iy0 - Y0
        win = (X0, Y0, max(0, X1 - X0), max(0, Y1 - Y0))
    assert abs(mux) <= 0.5 and abs(muy) <= 0.5
    J['mw'], J['mh'] = win[2], win[3]
    J['sx'], J['sy'] = sx, sy
    J['mux'], J['muy'] = mux, muy
    J['K'] = K
    J['amp'][:K] = amp
    J['var'][:K, 0] = var_pix[:, 0, 0]
    J['var'][:K, 1] = var_pix[:, 0, 1]
    J['var'][:K, 2] = var_pix[:, 1, 1]
    J['scale'] = 1.
    return sz, J, win


def point_job(px, py, psf_shape, mm):
    """Plan a PixelizedPSF point-source render on a modelMask
    (tractor/psf.py:137-190).  Returns the job or None (no overlap)."""
    H, W = psf_shape
    ix, iy = int(round(float(px))), int(round(float(py)))
    x0, y0 = ix - W // 2, iy - H // 2
    mx0, my0, mw, mh = mm
    if (mx0 >= x0 + W) or (mx0 + mw <= x0) or (my0 >= y0 + H) or (my0 + mh <= y0):
        return None
    J = np.zeros((), FFT_JOB)
    J['mux'], J['muy'] = px - ix, py - iy
    J['sx'], J['sy'] = x0 - mx0, y0 - my0
    J['mw'], J['mh'] = mw, mh
    J['scale'] = 1.
    return J


class PsfSpectra(object):
    """Device PSF spectra of one stamp, cached per FFT size (the reference's
    PixelizedPSF.fftcache, tractor/psf.py:268-280)."""
    def __init__(self, stamp, device='cuda'):
        stamp = np.ascontiguousarray(stamp, np.float32)
        self.shape = stamp.shape
        self.stamp = torch.from_numpy(stamp).to(device)
        self.device = device
        self.cache = {}

    def get(self, sz):
        if sz not in self.cache:
            if sz not in SUPPORTED_FFT_SIZES:
                raise ValueError('FFT size %d outside %s' % (sz, SUPPORTED_FFT_SIZES))
            ph, pw = self.shape
            out = torch.empty((sz // 2 + 1, sz, 2), dtype=torch.float32, device=self.device)
            lib = _lib.load()
            h = _lib.handle(torch.device(self.device).index or 0)
            import ctypes
            cx, cy = ctypes.c_int(), ctypes.c_int()
            check(lib.trh_psf_spectrum_batch(h.value, ptr(self.stamp), 1, ph, pw, sz, ptr(out),
                                             ctypes.byref(cx), ctypes.byref(cy), None),
                  'trh_psf_spectrum_batch')
            self.cache[sz] = (out, cx.value, cy.value)
        return self.cache[sz]


def jobs_to_device(jobs, device='cuda'):
    arr = np.ascontiguousarray(jobs)
    t = torch.from_numpy(arr.view(np.uint8).copy()).to(device)
    return t


def render_fft_jobs(sz, jobs, spectra, out=None, data=None, inverr=None, chi2=None, device='cuda'):
    """Run trh_render_fft_batch for records of one FFT size."""
    if len(jobs) == 0:
        return
    lib = _lib.load()
    h = _lib.handle(torch.device(device).index or 0)
    spec, _, _ = spectra.get(sz)
    dj = jobs_to_device(jobs, device)
    check(lib.trh_render_fft_batch(h.value, sz, ptr(dj), len(jobs), ptr(spec), ptr(out),
                                   ptr(data), ptr(inverr), ptr(chi2), None),
          'trh_render_fft_batch')


def render_point_jobs(jobs, spectra, out=None, data=None, inverr=None, chi2=None, device='cuda'):
    """Run trh_render_point_batch (PixelizedPSF stamps)."""
    if len(jobs) == 0:
        return
    lib = _lib.load()
    h = _lib.handle(torch.device(device).index or 0)
    ph, pw = spectra.shape
    dj = jobs_to_device(jobs, device)
    check(lib.trh_render_point_batch(h.value, ptr(dj), len(jobs), ptr(spectra.stamp), ph, pw,
                                     ptr(out), ptr(data), ptr(inverr), ptr(chi2), None),
          'trh_render_point_batch')
