"""Per-call device renders behind the object API: each function plans one job
(host bookkeeping), runs it through libtractor_hip and returns float32 pixels.

These are the MI355X replacements of the reference's per-source SWIG calls
(mp_fourier.gaussian_fourier_transform_zero_mean + irfftn + lanczos_shift_3f,
mix.c_gauss_2d_grid / c_gauss_2d_masked).  There is no CPU path: without the
library every call raises.
"""
import ctypes

import numpy as np
import torch

from . import _lib, render as R
from ._lib import FFT_JOB, MOG_JOB, TRH_STORE, check, ptr

_DEV = 'cuda'


def _h():
    return _lib.handle(0)


def _spectra(psf):
    sp = getattr(psf, '_trh_spectra', None)
    if sp is None:
        sp = R.PsfSpectra(psf.img, _DEV)
        psf._trh_spectra = sp
    return sp


def fft_render(amp, var_pix, px, py, psf, mm=None, nomask_extent=None):
    """FFT-path unit-flux render (tractor/galaxy.py:336-508) -> (x0, y0, patch)."""
    sz, J, win = R.galaxy_fft_job(np.asarray(amp), np.asarray(var_pix), px, py, psf.img.shape,
                                  mm=mm, nomask_extent=nomask_extent)
    X0, Y0, w, h = win
    out = torch.zeros(max(w * h, 1), dtype=torch.float32, device=_DEV)
    if w > 0 and h > 0:
        J['out_offset'] = 0
        J['out_stride'] = w
        J['mode'] = TRH_STORE
        R.render_fft_jobs(sz, np.array([J], FFT_JOB), _spectra(psf), out=out, device=_DEV)
    torch.cuda.synchronize()
    return X0, Y0, out[:w * h].cpu().numpy().reshape(h, w)


def point_render(psf, px, py, mm):
    """PixelizedPSF.getPointSourcePatch with a modelMask (tractor/psf.py:
    137-190) -> patch or None."""
    J = R.point_job(px, py, psf.img.shape, mm)
    if J is None:
        return None
    mx0, my0, mw, mh = mm
    out = torch.zeros(mw * mh, dtype=torch.float32, device=_DEV)
    J['out_offset'] = 0
    J['out_stride'] = mw
    J['mode'] = TRH_STORE
    R.render_point_jobs(np.array([J], FFT_JOB), _spectra(psf), out=out, device=_DEV)
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(mh, mw)


def lanczos_shift(img, dx, dy):
    """lanczos_shift_image (tractor/psf.py:26-58) on the device."""
    img = np.ascontiguousarray(img, np.float32)
    H, W = img.shape
    lib = _lib.load()
    din = torch.from_numpy(img.ravel().copy()).to(_DEV)
    dout = torch.empty_like(din)
    ddx = torch.tensor([float(dx)], dtype=torch.float64, device=_DEV)
    ddy = torch.tensor([float(dy)], dtype=torch.float64, device=_DEV)
    check(lib.trh_lanczos_shift_batch(_h().value, ptr(din), ptr(dout), 1, H, W, ptr(ddx), ptr(ddy),
                                      None), 'trh_lanczos_shift_batch')
    torch.cuda.synchronize()
    return dout.cpu().numpy().reshape(H, W)


def mog_render(amp, mean, var, x0, y0, w, h, fx=0., fy=0., mask=None, derivs=False,
               masked_rules=False):
    """Pixel-space MoG on the window [x0, x0+w) x [y0, y0+h), centred at
    mean + (fx, fy): c_gauss_2d_grid rules (tractor/mix.i:452-524) or, with
    masked_rules, c_gauss_2d_masked (tractor/gauss_masked.c) incl. the mask
    and derivatives.  Returns patch[, dx, dy] (float32)."""
    amp = np.asarray(amp, np.float64)
    mean = np.asarray(mean, np.float64)
    var = np.asarray(var, np.float64)
    K = len(amp)
    if w <= 0 or h <= 0:
        z = np.zeros((max(h, 0), max(w, 0)), np.float32)
        return (z, z.copy(), z.copy()) if derivs else z
    cx = mean[:, 0] + fx
    cy = mean[:, 1] + fy
    cxi = np.floor(cx).astype(np.int32)
    cyi = np.floor(cy).astype(np.int32)
    t = lambda a, dt=torch.float32: torch.from_numpy(np.ascontiguousarray(a)).to(_DEV, dt)
    d_amp = t(amp.astype(np.float32))
    d_cxi = t(cxi, torch.int32)
    d_cyi = t(cyi, torch.int32)
    d_cxf = t((cx - cxi).astype(np.float32))
    d_cyf = t((cy - cyi).astype(np.float32))
    v3 = np.stack([var[:, 0, 0], 0.5 * (var[:, 0, 1] + var[:, 1, 0]), var[:, 1, 1]], 1)
    d_var = t(v3.astype(np.float32))
    J = np.zeros((), MOG_JOB)
    J['out_offset'], J['out_stride'] = 0, w
    J['x0'], J['y0'], J['w'], J['h'] = x0, y0, w, h
    J['comp0'], J['K'] = 0, K
    J['cutoff_mode'] = 1 if masked_rules else 0
    J['mode'] = TRH_STORE
    J['want_derivs'] = 1 if derivs else 0
    J['scale'] = 1.
    d_mask = None
    if mask is not None:
        J['mask_offset'] = 0
        d_mask = torch.from_numpy(np.ascontiguousarray(mask, np.uint8)).to(_DEV)
    else:
        J['mask_offset'] = -1
    out = torch.zeros(w * h, dtype=torch.float32, device=_DEV)
    dx = torch.zeros_like(out) if derivs else None
    dy = torch.zeros_like(out) if derivs else None
    dj = R.jobs_to_device(np.array([J], MOG_JOB), _DEV)
    lib = _lib.load()
    check(lib.trh_render_mog_batch(_h().value, ptr(dj), 1, w * h, ptr(d_amp), ptr(d_cxi),
                                   ptr(d_cyi), ptr(d_cxf), ptr(d_cyf), ptr(d_var), ptr(d_mask),
                                   ptr(out), ptr(dx), ptr(dy), None), 'trh_render_mog_batch')
    torch.cuda.synchronize()
    o = out.cpu().numpy().reshape(h, w)
    if derivs:
        return o, dx.cpu().numpy().reshape(h, w), dy.cpu().numpy().reshape(h, w)
    return o


def mog_render_approx3(amp, mean, var, x0, x1, y0, y1, fx, fy, minval, xc, yc, minradius=3,
                       derivs=False):
    """c_gauss_2d_approx3 (tractor/approx3.c:1-465) on the device: ring-growth
    evaluation of the box [x0, x1) x [y0, y1).  Returns (result, xderiv,
    yderiv, (sx0, sx1, sy0, sy1)) with float64 arrays (derivatives None
    unless requested) and the box-relative bounds of the evaluated pixels."""
    from ._lib import MOG_APPROX_JOB, TRH_APPROX_MAXK, TRH_APPROX_MAX_WH
    amp = np.ascontiguousarray(amp, np.float64).ravel()
    mean = np.ascontiguousarray(mean, np.float64).reshape(-1, 2)
    var = np.ascontiguousarray(var, np.float64).reshape(-1, 2, 2)
    K = len(amp)
    W, H = int(x1 - x0), int(y1 - y0)
    if W <= 0 or H <= 0:
        raise ValueError('empty box')
    if K < 1 or K > TRH_APPROX_MAXK or len(mean) != K or len(var) != K:
        raise ValueError('approx3: 1..%d components with matching mean/var' % TRH_APPROX_MAXK)
    if W + H > TRH_APPROX_MAX_WH:
        raise ValueError('approx3: w + h = %d exceeds %d' % (W + H, TRH_APPROX_MAX_WH))
    J = np.zeros(1, MOG_APPROX_JOB)
    J['out_offset'], J['out_stride'] = 0, W
    J['x0'], J['y0'], J['w'], J['h'] = x0, y0, W, H
    J['comp0'], J['K'] = 0, K
    J['xc'], J['yc'] = xc, yc
    J['minradius'] = minradius
    J['want_derivs'] = 1 if derivs else 0
    J['fx'], J['fy'], J['minval'] = fx, fy, minval
    dj = R.jobs_to_device(J, _DEV)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(_DEV)
    d_amp, d_mean, d_var = t(amp), t(mean), t(var)
    out = torch.empty(W * H, dtype=torch.float64, device=_DEV)
    dxo = torch.empty_like(out) if derivs else None
    dyo = torch.empty_like(out) if derivs else None
    bbox = torch.empty(4, dtype=torch.int32, device=_DEV)
    lib = _lib.load()
    check(lib.trh_render_mog_approx3_batch(_h().value, ptr(dj), 1, W + H, ptr(d_amp), ptr(d_mean),
                                           ptr(d_var), ptr(out), ptr(dxo), ptr(dyo), ptr(bbox),
                                           None), 'trh_render_mog_approx3_batch')
    torch.cuda.synchronize()
    bb = tuple(int(v) for v in bbox.cpu().numpy())
    if bb[0] < 0:
        raise RuntimeError('approx3 job rejected by the device')
    res = out.cpu().numpy().reshape(H, W)
    if derivs:
        return res, dxo.cpu().numpy().reshape(H, W), dyo.cpu().numpy().reshape(H, W), bb
    return res, None, None, bb
