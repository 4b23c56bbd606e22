"""Single-call device renders: a pixel-space mixture through a one-request
RenderPlan, and the ring-growth (approx3) mixture evaluation.

These stand for the reference's per-call SWIG functions (mix.c_gauss_2d_grid
/ c_gauss_2d_masked, mix.c_gauss_2d_approx3).  There is no CPU path: without
the library every call raises.
"""
import numpy as np
import torch

from . import _lib, render as R
from ._lib import check, ptr

_DEV = 'cuda'


def _h():
    return _lib.handle(0)


def mog_render(amp, mean, var, x0, y0, w, h, fx=0., fy=0., mask=None, derivs=False,
               masked_rules=False):
    """Pixel-space MoG on the window [x0, x0+w) x [y0, y0+h), centred at
    mean + (fx, fy): c_gauss_2d_grid rules (tractor/mix.i:452-524) or, with
    masked_rules, c_gauss_2d_masked (tractor/gauss_masked.c) incl. the mask
    and derivatives.  Returns patch[, dx, dy] (float32)."""
    from .plan import RenderPlan
    plan = RenderPlan()
    s = plan.mog(amp, mean, var, x0, y0, w, h, fx, fy, mask=mask, derivs=derivs,
                 masked_rules=masked_rules)
    plan.run()
    return plan.result_derivs(s) if derivs else plan.result(s)


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
