"""FP64 NumPy restatement of the reference render-and-fit path (oracle).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.  Nothing in
``tractor_amd`` imports this module.

Arithmetic follows the reference: FP64 where the reference uses Python/NumPy
floats or C doubles, FP32 where the reference casts to float32 (the Lanczos
shift operates on float32 images with float32 taps, ``tractor/psf.py:50-56``,
``tractor/mp_fourier.i:267-352``; the masked MoG evaluator is FP32,
``tractor/gauss_masked.c``).
"""
import math

import numpy as np

__all__ = [
    'EXP_AMP', 'EXP_VAR', 'DEV_AMP', 'DEV_VAR', 'exp_mixture', 'dev_mixture',
    'galaxyshape_basis', 'ellipse_e_basis', 'ellipse_esoft_basis',
    'esoft_from_e', 'shear_mixture', 'get_overlapping_region',
    'mog_fourier_zero_mean', 'fft_size', 'pad_psf', 'psf_spectrum',
    'lanczos_taps3', 'correlate7f', 'lanczos_shift_image',
    'render_galaxy_fft', 'galaxy_patch_halfsize', 'mog_grid', 'mog_masked',
    'mog_convolve', 'render_galaxy_mog', 'mog_grid_approx3', 'mog_psf_point_patch', 'forced_unit_patch', 'forced_phot_lsq',
    'forced_fitstats', 'psfex_terms', 'psfex_at', 'em_fit_2d_reg', 'em_init_params',
    'pixpsf_point_patch', 'galaxy_param_derivs', 'pointsource_param_derivs',
    'column_scaled_lstsq', 'parameter_steps', 'accept_line_search', 'loop_should_stop',
    'SourceProblem', 'source_render', 'source_jacobian', 'source_update',
    'source_chisq', 'fit_source', 'ALPHAS', 'round_half_even', 'sersic_mixture',
    'sersic_fd_step', 'image_galaxy_unit_patch', 'image_model', 'image_chisq',
    'fit_image_lsqr',
]

# ---------------------------------------------------------------------------
# Mixture-of-Gaussians profiles (tractor/mixture_profiles.py:17-44).
# These tables are data the rendering must match exactly, so they are the
# same numbers as the reference.
# ---------------------------------------------------------------------------
_exp_amp = np.array([2.34853813e-03, 3.07995260e-02, 2.23364214e-01,
                     1.17949102e+00, 4.33873750e+00, 5.99820770e+00])
EXP_VAR = np.array([1.20078965e-03, 8.84526493e-03, 3.91463084e-02,
                    1.39976817e-01, 4.60962500e-01, 1.50159566e+00])
EXP_AMP = _exp_amp / np.sum(_exp_amp)

_dev_amp = np.array([4.26347652e-02, 2.40127183e-01, 6.85907632e-01,
                     1.51937350e+00, 2.83627243e+00, 4.46467501e+00,
                     5.72440830e+00, 5.60989349e+00])
_dev_var = np.array([2.23759216e-04, 1.00220099e-03, 4.18731126e-03,
                     1.69432589e-02, 6.84850479e-02, 2.87207080e-01,
                     1.33320254e+00, 8.40215071e+00])
_dev_core = 0.010233
# delta-function core prepended (mixture_profiles.py:32-38): K = 9
DEV_AMP = np.append(_dev_core, _dev_amp * (1. - _dev_core) / np.sum(_dev_amp))
DEV_VAR = np.append(0., _dev_var)


def _iso_var(v):
    var = np.zeros((len(v), 2, 2))
    var[:, 0, 0] = v
    var[:, 1, 1] = v
    return var


def exp_mixture():
    """ExpGalaxy profile, normalised (galaxy.py:727-731)."""
    amp = EXP_AMP / np.sum(EXP_AMP)
    return amp.copy(), _iso_var(EXP_VAR)


def dev_mixture():
    """DevGalaxy profile incl. the delta core, normalised (galaxy.py:746-750)."""
    amp = DEV_AMP / np.sum(DEV_AMP)
    return amp.copy(), _iso_var(DEV_VAR)


def round_half_even(x):
    """Python 3 ``round`` / ``np.round`` semantics (round half to even)."""
    return int(np.round(x))


# ---------------------------------------------------------------------------
# Shapes (tractor/galaxy.py:66-88, tractor/ellipses.py:126-155,214-255)
# ---------------------------------------------------------------------------
def galaxyshape_basis(re, ab, phi):
    """GalaxyShape.getRaDecBasis (galaxy.py:74-88): r_e -> degrees."""
    phi = np.deg2rad(90 - phi)
    re_deg = max(1. / 30, re) / 3600.
    cp = np.cos(phi)
    sp = np.sin(phi)
    return re_deg * np.array([[cp, sp * ab], [-sp, cp * ab]])


def ellipse_e_basis(re, e1, e2):
    """EllipseE.getRaDecBasis (ellipses.py:126-155)."""
    theta = math.atan2(e2, e1) / 2.
    ct = math.cos(theta)
    st = math.sin(theta)
    e = math.hypot(e1, e2)
    maxab = 1000.
    if e >= 1.:
        ab = maxab
    else:
        ab = min(maxab, (1. + e) / (1. - e))
    r_deg = re / 3600.
    return r_deg * np.array([[ct / ab, st], [-st / ab, ct]])


def ellipse_esoft_basis(logre, ee1, ee2):
    """EllipseESoft: re=exp(logre), e=1-exp(-|ee|), theta=atan2(ee2,ee1)/2
    (ellipses.py:232-255), then the EllipseE basis (ellipses.py:126-155)."""
    re = math.exp(logre)
    ee = math.hypot(ee1, ee2)
    e = 1. - math.exp(-ee)
    theta = math.atan2(ee2, ee1) / 2.
    ct = math.cos(theta)
    st = math.sin(theta)
    maxab = 1000.
    if e >= 1.:
        ab = maxab
    else:
        ab = min(maxab, (1. + e) / (1. - e))
    r_deg = re / 3600.
    return r_deg * np.array([[ct / ab, st], [-st / ab, ct]])


def esoft_from_e(re, e1, e2, maxe=0.999999):
    """EllipseESoft.fromEllipseE (ellipses.py:199-203)."""
    e = min(math.hypot(e1, e2), maxe)
    esoft = -math.log(1. - e)
    if e == 0:
        return (math.log(re), 0., 0.)
    return (math.log(re), e1 / e * esoft, e2 / e * esoft)


def shear_mixture(var, cdinv, basis):
    """HoggGalaxy._getShearedProfile (galaxy.py:659-669) with
    MixtureOfGaussians.apply_shear (mixture_profiles.py:146-155):
    var' = T var T^T with T = cdinv . G."""
    T = np.dot(cdinv, basis)
    return T @ var @ T.T


# ---------------------------------------------------------------------------
# get_overlapping_region (tractor/patch.pyx:20-66): inclusive ranges.
# ---------------------------------------------------------------------------
def get_overlapping_region(xlo, xhi, xmin, xmax):
    if xlo > xmax or xhi < xmin or xlo > xhi or xmin > xmax:
        return ([], [])
    xloclamp = max(xlo, xmin)
    Xlo = xloclamp - xlo
    xhiclamp = min(xhi, xmax)
    Xhi = Xlo + (xhiclamp - xloclamp)
    return (slice(xloclamp, xhiclamp + 1), slice(Xlo, Xhi + 1))


# ---------------------------------------------------------------------------
# MoG Fourier transform (mp_fourier.i:431-478; mixture_profiles.py:174-218)
# ---------------------------------------------------------------------------
def mog_fourier_zero_mean(amp, var, v, w):
    """F[j,i] = sum_k amp_k exp(-2 pi^2 (a v_i^2 + 2 b v_i w_j + d w_j^2));
    a delta component (a=b=d=0) contributes amp_k (mp_fourier.i:465-469)."""
    v = np.asarray(v, np.float64)[np.newaxis, :]
    w = np.asarray(w, np.float64)[:, np.newaxis]
    out = np.zeros((w.shape[0], v.shape[1]), np.float64)
    negtwopisq = -2. * np.pi * np.pi
    for k in range(len(amp)):
        a = var[k, 0, 0]
        b = var[k, 0, 1]
        d = var[k, 1, 1]
        if a == 0. and b == 0. and d == 0.:
            out += amp[k]
            continue
        out += amp[k] * np.exp(negtwopisq * (a * v * v + 2. * b * v * w + d * w * w))
    return out


# ---------------------------------------------------------------------------
# Pixelized PSF spectrum (psf.py:209-280)
# ---------------------------------------------------------------------------
def fft_size(radius):
    """PixelizedPSF.getFourierTransformSize (psf.py:209-212)."""
    return 2 ** int(np.ceil(np.log2(radius * 2.)))


def pad_psf(img, H, W):
    """PixelizedPSF._padInImage (psf.py:214-247) -> (pad, cx, cy)."""
    ph, pw = img.shape
    subimg = img
    if H >= ph:
        y0 = (H - ph) // 2
        cy = y0 + ph // 2
    else:
        y0 = 0
        cut = (ph - H) // 2
        subimg = subimg[cut:cut + H, :]
        cy = ph // 2 - cut
    if W >= pw:
        x0 = (W - pw) // 2
        cx = x0 + pw // 2
    else:
        x0 = 0
        cut = (pw - W) // 2
        subimg = subimg[:, cut:cut + W]
        cx = pw // 2 - cut
    sh, sw = subimg.shape
    pad = np.zeros((H, W), img.dtype)
    pad[y0:y0 + sh, x0:x0 + sw] = subimg
    return pad, cx, cy


def psf_spectrum(img, radius):
    """PixelizedPSF.getFourierTransform (psf.py:249-280).
    Returns (P complex64 (sz, sz/2+1), (cx, cy), (sz, sz), (v, w))."""
    img = np.asarray(img, np.float32)
    sz = fft_size(radius)
    pad, cx, cy = pad_psf(img, sz, sz)
    # The reference ran numpy < 2, whose rfft2 upcasts float32 input to
    # float64; do the same, then cast to complex64 (psf.py:273-274).
    P = np.fft.rfft2(pad.astype(np.float64)).astype(np.complex64)
    v = np.fft.rfftfreq(sz)
    w = np.fft.fftfreq(sz)
    return P, (cx, cy), (sz, sz), (v, w)


# ---------------------------------------------------------------------------
# Lanczos-3 shift (mp_fourier.i:267-429; psf.py:26-58)
# ---------------------------------------------------------------------------
def lanczos_taps3(dx):
    """lanczos_filter_3 (mp_fourier.i:355-402): 7 FP64 taps, the tap beyond
    +-3 zeroed (last one for dx>0, first for dx<0), normalised to sum 1."""
    assert -0.5 <= dx <= 0.5
    filt = np.zeros(7)
    if dx == 0.0:
        filt[3] = 1.0
        return filt
    idx = range(0, 6) if dx > 0 else range(1, 7)
    s = 0.
    for i in idx:
        pix = math.pi * (i - 3.0 + dx)
        val = 3 * math.sin(pix) * math.sin(pix / 3.0) / (pix * pix)
        filt[i] = val
        s += val
    for i in idx:
        filt[i] /= s
    return filt


def _corr7_rows_f32(img, f):
    """One pass of correlate7f_inout (mp_fourier.i:290-320): for every row,
    out[x] = sum_{k=0..6} f[k] in[x-3+k] with zero padding, FP32 arithmetic,
    accumulated in k order."""
    img = np.asarray(img, np.float32)
    H, W = img.shape
    f = np.asarray(f, np.float64).astype(np.float32)
    pad = np.zeros((H, W + 6), np.float32)
    pad[:, 3:3 + W] = img
    acc = np.zeros((H, W), np.float32)
    for k in range(7):
        acc = (acc + f[k] * pad[:, k:k + W]).astype(np.float32)
    return acc


def correlate7f(img, fx, fy):
    """correlate7f_inout (mp_fourier.i:267-352): filter rows with fx, then
    columns with fy."""
    work = _corr7_rows_f32(img, fx)
    return _corr7_rows_f32(work.T, fy).T.copy()


def lanczos_shift_image(img, dx, dy):
    """lanczos_shift_image (psf.py:26-58) -> lanczos_shift_3f
    (mp_fourier.i:405-429).  out(x,y) ~= in(x-dx, y-dy).  The reference's
    scipy fallback for W<=8 or H<=8 (psf.py:36-49) uses the same zero-padded
    7-tap correlation with the same normalised taps."""
    return correlate7f(np.asarray(img, np.float32), lanczos_taps3(dx), lanczos_taps3(dy))


# ---------------------------------------------------------------------------
# MoG pixel evaluators (mix.i:452-524; gauss_masked.c; mixture_profiles.py)
# ---------------------------------------------------------------------------
def mog_grid(x0, x1, y0, y1, fx, fy, amp, mean, var):
    """c_gauss_2d_grid (mix.i:452-524), FP64: result[iy,ix] =
    sum_k amp_k/(2 pi sqrt(det)) exp(-1/2 d), skipping d >= 100."""
    NX = x1 - x0
    NY = y1 - y0
    result = np.zeros((NY, NX))
    xx = np.arange(x0, x1, dtype=np.float64)[np.newaxis, :]
    yy = np.arange(y0, y1, dtype=np.float64)[:, np.newaxis]
    tpd = (2. * np.pi) ** 2
    for k in range(len(amp)):
        V = var[k]
        det = V[0, 0] * V[1, 1] - V[0, 1] * V[1, 0]
        I0 = V[1, 1] / det
        I1 = -(V[0, 1] + V[1, 0]) / det
        I2 = V[0, 0] / det
        scale = amp[k] / np.sqrt(tpd * det)
        dx = xx - fx - mean[k, 0]
        dy = yy - fy - mean[k, 1]
        dsq = I0 * dx * dx + I1 * dx * dy + I2 * dy * dy
        g = np.where(dsq >= 100, 0., scale * np.exp(-0.5 * np.minimum(dsq, 100.)))
        result += g
    return result


def mog_masked(x0, y0, mask, fx, fy, amp, mean, var, derivs=False):
    """c_gauss_2d_masked (gauss_masked.c:1-159) with eval_all_dxy_f
    (mix.i:85-142), FP32: the -1/2 and the cross-term factor 2 are folded
    into icov; components with exponent < -30 are skipped; non-finite
    variances/scales zero the component; derivatives are with respect to the
    mean."""
    f32 = np.float32
    mask = np.asarray(mask, bool)
    H, W = mask.shape
    amp = np.asarray(amp, f32)
    mean = np.asarray(mean, f32)
    var = np.asarray(var, f32)
    fx = f32(fx)
    fy = f32(fy)
    result = np.zeros((H, W), f32)
    xd = np.zeros((H, W), f32)
    yd = np.zeros((H, W), f32)
    tpd = f32((2. * np.pi) ** 2)
    xx = (np.arange(W, dtype=np.int64) + x0).astype(f32)[np.newaxis, :] - fx
    yy = (np.arange(H, dtype=np.int64) + y0).astype(f32)[:, np.newaxis] - fy
    with np.errstate(all='ignore'):
        for k in range(len(amp)):
            V0 = var[k, 0, 0]
            V1 = f32((var[k, 0, 1] + var[k, 1, 0]) * f32(0.5))
            V2 = var[k, 1, 1]
            det = f32(V0 * V2 - V1 * V1)
            isc = f32(f32(-0.5) / det)
            I0 = f32(V2 * isc)
            I1 = f32(f32(-V1 * isc) * f32(2.0))
            I2 = f32(V0 * isc)
            scale = f32(amp[k] / np.sqrt(f32(tpd * det)))
            if not all(np.isfinite([V0, V1, V2, I0, I1, I2, scale])):
                continue
            if scale == 0:
                continue
            dx = (xx - mean[k, 0]).astype(f32)
            dy = (yy - mean[k, 1]).astype(f32)
            dsq = (I0 * dx * dx + I1 * dx * dy + I2 * dy * dy).astype(f32)
            ok = (dsq >= f32(-30.)) & np.isfinite(dsq)
            G = np.where(ok, scale * np.exp(np.where(ok, dsq, 0)).astype(f32), 0).astype(f32)
            result += G
            if derivs:
                xd += (-G * (f32(2.) * I0 * dx + I1 * dy)).astype(f32)
                yd += (-G * (f32(2.) * I2 * dy + I1 * dx)).astype(f32)
    result[~mask] = 0
    if derivs:
        xd[~mask] = 0
        yd[~mask] = 0
        return result, xd, yd
    return result


def _approx3_prep(amp, var, minval):
    """Per-component constants of c_gauss_2d_approx3 (approx3.c:96-126):
    icov with -1/2 and the cross factor 2 folded in, scale = amp/(2 pi sqrt
    det), and the exponent floor maxD = log(minval 2 pi sqrt(det) / sum amp),
    clamped at -100 (or -100 when minval <= 0)."""
    tpd = (2. * np.pi) ** 2
    ampsum = float(np.sum(amp))
    comps = []
    for k in range(len(amp)):
        V0 = var[k][0][0]
        V1 = (var[k][0][1] + var[k][1][0]) * 0.5
        V2 = var[k][1][1]
        det = V0 * V2 - V1 * V1
        isc = -0.5 / det
        icov = (V2 * isc, -V1 * isc * 2.0, V0 * isc)
        scale = amp[k] / math.sqrt(tpd * det)
        if minval > 0.:
            md = math.log(minval * math.sqrt(tpd * det) / ampsum)
            md = max(md, -100.)
        else:
            md = -100.
        comps.append((icov, scale, md))
    return comps


def _approx3_eval(comps, mean, x, y, cutoff):
    """eval_all (mix.i:31-42, exponent floor -100) when cutoff is False,
    eval_all_dxy (mix.i:44-82, per-component floor maxD, non-finite skipped)
    when True.  Returns (value, d/dmean_x, d/dmean_y)."""
    r = dxs = dys = 0.
    for k, (I, scale, md) in enumerate(comps):
        dx = x - mean[k][0]
        dy = y - mean[k][1]
        dsq = I[0] * dx * dx + I[1] * dx * dy + I[2] * dy * dy
        if cutoff:
            if not (dsq >= md) or not math.isfinite(dsq):
                continue
        elif dsq < -100:
            continue
        G = scale * math.exp(dsq)
        r += G
        dxs += -G * (2. * I[0] * dx + I[1] * dy)
        dys += -G * (2. * I[2] * dy + I[1] * dx)
    return r, dxs, dys


def mog_grid_approx3(x0, x1, y0, y1, fx, fy, minval, amp, mean, var, xc, yc, minradius=3):
    """c_gauss_2d_approx3 (approx3.c:1-465), FP64, pure Python (small cases).

    Evaluation grows outward in square rings around the start pixel.  Each
    ring has four sides: top row (y = yc+R), bottom row (y = yc-R), left
    column (x = xc-R) and right column (x = xc+R); a side pixel is evaluated
    only if it was marked by the previous ring.  An evaluated pixel whose
    value is >= minval, or any pixel while R <= minradius, marks its outward
    neighbours for the next ring: a top-row pixel marks the three pixels
    above it and, at the ring's left/right end, two rows of the next left /
    right column; bottom mirrors top; a left-column pixel marks three rows of
    the next left column; right mirrors left.  Growth stops at the first ring
    with no marked pixel.  If (xc, yc) lies outside the box, the start pixel
    is the maximum of eval_all along the nearest edge(s) (first maximum in
    scan order; box centre when every value is 0).

    Returns (result, xderiv, yderiv, (sx0, sx1, sy0, sy1)) with the
    half-open bounding box of evaluated pixels (box coordinates).
    """
    W = x1 - x0
    H = y1 - y0
    comps = _approx3_prep(amp, var, minval)
    res = np.zeros((H, W))
    xd = np.zeros((H, W))
    yd = np.zeros((H, W))
    if xc < x0 or yc < y0 or xc >= x1 or yc >= y1:
        best = 0.
        nx, ny = (x0 + x1) // 2, (y0 + y1) // 2
        cand = []
        if xc < x0 or xc >= x1:
            ex = x0 if xc < x0 else x1 - 1
            cand += [(ex, iy) for iy in range(y0, y1)]
        if yc < y0 or yc >= y1:
            ey = y0 if yc < y0 else y1 - 1
            cand += [(ix, ey) for ix in range(x0, x1)]
        for (ix, iy) in cand:
            v = _approx3_eval(comps, mean, ix - fx, iy - fy, False)[0]
            if v > best:
                best, nx, ny = v, ix, iy
        xc, yc = nx, ny

    def ev(x, y):
        r, a, b = _approx3_eval(comps, mean, x - fx, y - fy, True)
        res[y - y0, x - x0] = r
        xd[y - y0, x - x0] = a
        yd[y - y0, x - x0] = b
        return r

    # marks for the next ring, in image coordinates: columns for top/bottom
    # rows, rows for left/right columns; kept inside the box.
    def inx(v):
        return x0 <= v < x1

    def iny(v):
        return y0 <= v < y1

    nT, nB, nL, nR = set(), set(), set(), set()
    if yc > y0:
        nB |= {v for v in (xc - 1, xc, xc + 1) if inx(v)}
    if yc < y1 - 1:
        nT |= {v for v in (xc - 1, xc, xc + 1) if inx(v)}
    if xc > x0 and iny(yc):
        nL.add(yc)
    if xc < x1 - 1 and iny(yc):
        nR.add(yc)
    ev(xc, yc)
    bx0 = bx1 = xc - x0
    by0 = by1 = yc - y0
    R = 0
    while True:
        R += 1
        T, B, L, Rt = nT, nB, nL, nR
        nT, nB, nL, nR = set(), set(), set(), set()
        anything = False
        yy = yc + R
        if yy < y1 and T:
            anything = True
            by1 = yy - y0
            for xx in sorted(T):
                r = ev(xx, yy)
                if R > minradius and r < minval:
                    continue
                if xx == xc - R and xx > x0:
                    nL |= {v for v in (yy, yy - 1) if iny(v)}
                if xx == xc + R and xx < x1 - 1:
                    nR |= {v for v in (yy, yy - 1) if iny(v)}
                if yy < y1 - 1:
                    nT |= {v for v in (xx - 1, xx, xx + 1) if inx(v)}
        yy = yc - R
        if yy >= y0 and B:
            anything = True
            by0 = yy - y0
            for xx in sorted(B):
                r = ev(xx, yy)
                if R > minradius and r < minval:
                    continue
                if xx == xc - R and xx > x0:
                    nL |= {v for v in (yy, yy + 1) if iny(v)}
                if xx == xc + R and xx < x1 - 1:
                    nR |= {v for v in (yy, yy + 1) if iny(v)}
                if yy > y0:
                    nB |= {v for v in (xx - 1, xx, xx + 1) if inx(v)}
        xx = xc - R
        if xx >= x0 and L:
            anything = True
            bx0 = xx - x0
            for yy in sorted(L):
                r = ev(xx, yy)
                if R > minradius and r < minval:
                    continue
                if xx > x0:
                    nL |= {v for v in (yy - 1, yy, yy + 1) if iny(v)}
        xx = xc + R
        if xx < x1 and Rt:
            anything = True
            bx1 = xx - x0
            for yy in sorted(Rt):
                r = ev(xx, yy)
                if R > minradius and r < minval:
                    continue
                if xx < x1 - 1:
                    nR |= {v for v in (yy - 1, yy, yy + 1) if iny(v)}
        if not anything:
            break
    return res, xd, yd, (bx0, bx1 + 1, by0, by1 + 1)


def mog_convolve(amp, mean, var, pamp, pmean, pvar):
    """MixtureOfGaussians.convolve (mixture_profiles.py:158-172): PSF
    component index is the outer loop."""
    K, Kp = len(amp), len(pamp)
    namp = np.zeros(K * Kp)
    nmean = np.zeros((K * Kp, 2))
    nvar = np.zeros((K * Kp, 2, 2))
    for k in range(Kp):
        sl = slice(k * K, (k + 1) * K)
        namp[sl] = amp * pamp[k]
        nmean[sl] = mean + pmean[k]
        nvar[sl] = var + pvar[k]
    return namp, nmean, nvar


# ---------------------------------------------------------------------------
# Galaxy rendering (galaxy.py:256-519, 640-691)
# ---------------------------------------------------------------------------
def galaxy_patch_halfsize(radius_px, px, py, img_shape, psf_radius):
    """HoggGalaxy._getUnitFluxPatchSize (galaxy.py:678-691) with
    NullWCS pixscale 1: radius_px = nre * re."""
    halfsize = max(1., radius_px)
    h, w = img_shape
    imsize = max(px, w - 1 - px, py, h - 1 - py) + 1
    halfsize = min(halfsize, imsize)
    halfsize = max(halfsize, psf_radius)
    return int(np.ceil(halfsize))


def render_galaxy_mog(amp, var_pix, px, py, pamp, pmean, pvar, x0, x1, y0, y1,
                      mask=None):
    """run_mog branch of _realGetUnitFluxModelPatch (galaxy.py:307-334):
    affine profile (mean = (px,py)) convolved with the PSF MoG, evaluated on
    the grid (minval=0 -> mixture_to_patch -> evaluate_grid) or masked."""
    mean = np.zeros((len(amp), 2)) + np.array([px, py])
    camp, cmean, cvar = mog_convolve(amp, mean, var_pix, pamp, pmean, pvar)
    if mask is not None:
        return mog_masked(x0, y0, mask, 0., 0., camp, cmean, cvar)
    return mog_grid(x0, x1, y0, y1, 0., 0., camp, cmean, cvar)


def render_galaxy_fft(amp, var_pix, px, py, psf_img, mm=None, nomask_extent=None,
                      hybrid_mog=None, force_sourceOut=False, radius_px=None,
                      inner_real_nsigma=3., outer_real_nsigma=4., mask=None):
    """FFT branch of ProfileGalaxy._realGetUnitFluxModelPatch
    (galaxy.py:336-519) for a PixelizedPSF (or HybridPixelizedPSF when
    ``hybrid_mog`` = (pamp, pmean, pvar) is given).

    mm: (x0, y0, w, h) modelMask, or None with ``nomask_extent`` =
        (x0, x1, y0, y1, halfsize) from the halfsize rule (galaxy.py:273-290).
    mask: the modelMask's boolean pixel map, if it has one (only the MoG
        branch reads it, galaxy.py:326-329).
    Returns (x0, y0, patch float32) or None.
    """
    psf_img = np.asarray(psf_img, np.float32)
    psfh, psfw = psf_img.shape
    if mm is not None:
        x0, y0, mw, mh = mm
        x1 = x0 + mw
        y1 = y0 + mh
        halfsize = max(mh / 2., mw / 2.)
        halfsize = max(halfsize, max(max(1 + px - x0, 1 + x1 - px),
                                     max(1 + py - y0, 1 + y1 - py)))
        halfsize = max(halfsize, max(psfw / 2., psfh / 2.))
        halfsize = int(np.ceil(halfsize))
        sourceOut = (px < x0 or px > x1 - 1 or py < y0 or py > y1 - 1)
        if sourceOut:
            neardx, neardy = 0., 0.
            if px < x0:
                neardx = x0 - px
            if px > x1:
                neardx = px - x1
            if py < y0:
                neardy = y0 - py
            if py > y1:
                neardy = py - y1
            nearest = np.hypot(neardx, neardy)
            if radius_px is not None and nearest > radius_px and not force_sourceOut:
                return None
            if hybrid_mog is not None:
                pamp, pmean, pvar = hybrid_mog
                return (x0, y0, render_galaxy_mog(amp, var_pix, px, py, pamp, pmean, pvar,
                                                  x0, x1, y0, y1, mask=mask))
            raise RuntimeError('sourceOut, not hybrid')
    else:
        x0, x1, y0, y1, halfsize = nomask_extent
    P, (cx, cy), (pH, pW), (v, w) = psf_spectrum(psf_img, halfsize)
    dx = px - cx
    dy = py - cy
    if mm is not None:
        ix0, iy0 = x0, y0
        mux = dx - ix0
        muy = dy - iy0
        sx = round_half_even(mux)
        sy = round_half_even(muy)
        mux -= sx
        muy -= sy
    else:
        ix0 = round_half_even(dx)
        iy0 = round_half_even(dy)
        mux = dx - ix0
        muy = dy - iy0
    assert abs(mux) <= 0.5 and abs(muy) <= 0.5

    fftamp = np.asarray(amp, np.float64)
    fftvar = var_pix
    mogmix = None
    if hybrid_mog is not None:
        vv = var_pix[:, 0, 0] + var_pix[:, 1, 1]
        n1, n2 = inner_real_nsigma, outer_real_nsigma
        IM = ((pW / 2) ** 2 < (n2 ** 2 * vv))
        IF = ((pW / 2) ** 2 > (n1 ** 2 * vv))
        ramp = np.any(IM * IF)
        if np.any(IM):
            amps = np.array(amp)[IM]
            if ramp:
                ns = (pW / 2) / np.maximum(1e-6, np.sqrt(vv))
                amps = amps * np.minimum(1., (n2 - ns[IM]) / (n2 - n1))
            mogmix = (amps, var_pix[IM])
        if np.any(IF):
            amps = np.array(amp)[IF]
            if ramp:
                ns = (pW / 2) / np.maximum(1e-6, np.sqrt(vv))
                amps = amps * np.minimum(1., (ns[IF] - n1) / (n2 - n1))
            fftamp, fftvar = amps, var_pix[IF]
        else:
            fftamp = None

    if fftamp is not None:
        Fsum = mog_fourier_zero_mean(fftamp, fftvar, v, w)
        G = np.fft.irfftn(Fsum * P)
        G = G.astype(np.float32)
        if mux != 0.0 or muy != 0.0:
            G = lanczos_shift_image(G, mux, muy)
    else:
        G = np.zeros((pH, pW), np.float32)

    if mm is not None:
        gh, gw = G.shape
        if sx != 0 or sy != 0:
            yi, yo = get_overlapping_region(-sy, -sy + mh - 1, 0, gh - 1)
            xi, xo = get_overlapping_region(-sx, -sx + mw - 1, 0, gw - 1)
            shG = np.zeros((mh, mw), G.dtype)
            if yi != [] and xi != []:
                shG[yo, xo] = G[yi, xi]
            G = shG
        if gh > mh or gw > mw:
            G = G[:mh, :mw]
    else:
        if x0 > ix0:
            G = G[:, x0 - ix0:]
            ix0 = x0
        if y0 > iy0:
            G = G[y0 - iy0:, :]
            iy0 = y0
        gh, gw = G.shape
        if gw + ix0 > x1:
            G = G[:, :x1 - ix0]
        if gh + iy0 > y1:
            G = G[:y1 - iy0, :]
    if mogmix is not None:
        pamp, pmean, pvar = hybrid_mog
        gh, gw = G.shape
        mamp, mvar = mogmix
        # run_mog on the modelMask, or on ModelMask(ix0, iy0, gw, gh) without
        # one (galaxy.py:510-516): the masked FP32 evaluator only when the
        # mask has a pixel map, else the FP64 grid (galaxy.py:326-332); the
        # sum lands in the float32 G (G += patch)
        m = render_galaxy_mog(mamp, mvar, px, py, pamp, pmean, pvar,
                              ix0, ix0 + gw, iy0, iy0 + gh,
                              mask=mask if mm is not None else None)
        G = (G + m).astype(np.float32)
    return (ix0, iy0, G)


# ---------------------------------------------------------------------------
# Point sources (pointsource.py:102-125; psf.py:137-207, 458-528, 898-929)
# ---------------------------------------------------------------------------
def mog_psf_point_patch(pamp, pmean, pvar, px, py, x0, x1, y0, y1):
    """MoG PSF point-source patch with minval=0: exact grid evaluation of the
    PSF mixture shifted to (px, py) (psf.py:516-519, 925-928)."""
    return mog_grid(x0, x1, y0, y1, px, py, pamp, pmean, pvar)


def pixpsf_point_patch(psf_img, px, py, mm=None, derivs=False):
    """PixelizedPSF.getPointSourcePatch (psf.py:137-207), sampling 1.
    Returns (x0, y0, patch[, dx, dy]) or None."""
    img = np.asarray(psf_img, np.float32)
    H, W = img.shape
    ix = round_half_even(float(px))
    iy = round_half_even(float(py))
    dx = px - ix
    dy = py - iy
    x0 = ix - W // 2
    y0 = iy - H // 2
    if mm is None:
        return (x0, y0, lanczos_shift_image(img, dx, dy))
    mx0, my0, mw, mh = mm
    if (mx0 >= x0 + W) or (mx0 + mw <= x0) or (my0 >= y0 + H) or (my0 + mh <= y0):
        return None
    L = 3
    pad = np.zeros((mh + 2 * L, mw + 2 * L), np.float32)
    yi, yo = get_overlapping_region(my0 - y0 - L, my0 - y0 + mh - 1 + L, 0, H - 1)
    xi, xo = get_overlapping_region(mx0 - x0 - L, mx0 - x0 + mw - 1 + L, 0, W - 1)
    pad[yo, xo] = img[yi, xi]
    pad = lanczos_shift_image(pad, dx, dy)
    p = pad[L:-L, L:-L].copy()
    if not derivs:
        return (mx0, my0, p)
    ddx = np.zeros_like(p)
    ddx[1:-1, 1:-1] = -(p[1:-1, 2:] - p[1:-1, :-2]) / 2.
    ddy = np.zeros_like(p)
    ddy[1:-1, 1:-1] = -(p[2:, 1:-1] - p[:-2, 1:-1]) / 2.
    return (mx0, my0, p, ddx, ddy)


# ---------------------------------------------------------------------------
# Finite-difference derivatives (galaxy.py:137-223; pointsource.py:130-241;
# ducks.py:413-423)
# ---------------------------------------------------------------------------
def _central_diff(p0):
    dx = np.zeros_like(p0)
    dx[1:-1, 1:-1] = (p0[1:-1, :-2] - p0[1:-1, 2:]) / 2.
    dy = np.zeros_like(p0)
    dy[1:-1, 1:-1] = (p0[:-2, 1:-1] - p0[2:, 1:-1]) / 2.
    return dx, dy


def galaxy_param_derivs(render, counts, cdi, shape_steps, flux_dcounts=1.):
    """Galaxy.getParamDerivatives (galaxy.py:137-223) for (pos[2], flux[1],
    shape[n]) with every parameter thawed.  ``render(k)`` returns the unit
    flux patch (float32) with shape parameter k stepped (k=None: current).
    Returns (patch0, [d/dpos_x, d/dpos_y, d/dflux, d/dshape_0, ...])."""
    p0 = render(None)
    dxp, dyp = _central_diff(p0)
    derivs = []
    for i in range(2):
        derivs.append(((dxp * cdi[0, i] + dyp * cdi[1, i]) * counts).astype(np.float32))
    derivs.append((p0 * flux_dcounts).astype(np.float32))
    for k, step in enumerate(shape_steps):
        pk = render(k)
        derivs.append(((pk - p0) * (counts / step)).astype(np.float32))
    return p0, derivs


def pointsource_param_derivs(p0, pdx, pdy, counts, cdi, flux_dcounts=1.):
    """PointSource.getParamDerivatives with fastPosDerivs and PSF-provided
    dx, dy (pointsource.py:204-206, 228-241; ducks.py:413-423)."""
    derivs = []
    for i in range(2):
        derivs.append(((pdx * cdi[0, i] + pdy * cdi[1, i]) * counts).astype(np.float32))
    derivs.append((p0 * flux_dcounts).astype(np.float32))
    return derivs


# ---------------------------------------------------------------------------
# Update direction and line search
# ---------------------------------------------------------------------------
def column_scaled_lstsq(A, b, use_lsqr=False, damp=0.):
    """Column-scaled linear least squares (lsqr_optimizer.py:470-488,
    584-591, 629-634; gpu_optimizer.py:758-792): scale = ||col||_2, solve on
    A/scale, unscale.  ``use_lsqr`` reproduces scipy's LSQR (default
    tolerances) as LsqrOptimizer does; otherwise a dense FP64 lstsq (the
    exact solution LSQR converges to)."""
    A = np.asarray(A, np.float64)
    b = np.asarray(b, np.float64)
    colscales = np.sqrt((A ** 2).sum(axis=0))
    live = colscales > 0
    X = np.zeros(A.shape[1])
    if not np.any(live):
        return None
    As = A[:, live] / colscales[live]
    if use_lsqr:
        from scipy.sparse.linalg import lsqr
        xs = lsqr(As, b, damp=damp)[0]
    else:
        xs = np.linalg.lstsq(As, b, rcond=None)[0]
    X[live] = xs / colscales[live]
    return X


ALPHAS = np.append(2. ** np.arange(-10, 1), [np.sqrt(2.), 2.])


def parameter_steps(p0, X, maxstep, lowers, uppers, alphas=None, tiny_alpha=1e-8,
                    flags=None):
    """ConstrainedOptimizer.getParameterSteps (constrained_optimizer.py:84-159).
    Returns [(alpha, p, step_limit, hit_limit), ...].  ``flags`` (a dict, if
    given) gets 'step_limited' = True wherever the reference sets
    self.stepLimited (line 121: any max-step clamp, including one whose alpha
    then falls below tiny_alpha and is not returned)."""
    if alphas is None:
        alphas = ALPHAS
    X = np.asarray(X, np.float64)
    results = []
    for alpha in alphas:
        do_break = False
        step_limit = False
        for i, mx in enumerate(maxstep):
            if mx is None:
                continue
            step = alpha * abs(X[i])
            if step > mx:
                step_limit = True
                do_break = True
                alpha = mx / abs(X[i])
                if flags is not None:
                    flags['step_limited'] = True
        hit_limit = False
        for i, (ps, l, u) in enumerate(zip(p0, lowers, uppers)):
            px = ps + alpha * X[i]
            if l is not None and px < l:
                alpha = (l - ps) / X[i]
                px = l
                hit_limit = True
            if u is not None and px > u:
                alpha = (u - ps) / X[i]
                px = u
                hit_limit = True
        if alpha < tiny_alpha:
            break
        p = []
        for ps, s, l, u in zip(p0, X, lowers, uppers):
            px = ps + alpha * s
            if l is not None and px < l:
                px = l
            if u is not None and px > u:
                px = u
            p.append(px)
        results.append((alpha, p, step_limit, hit_limit))
        if do_break:
            break
    return results


def accept_line_search(lnp0, steps, lnps):
    """ConstrainedOptimizer.tryUpdates acceptance (constrained_optimizer.py:
    36-82): walk the steps in order; stop at a non-finite lnp or one more than
    1 below the best; keep the best improving step.
    Returns (index or None, dlnp, alpha, last_step_hit_limit)."""
    best = lnp0
    ibest = None
    for i, ((alpha, p, sl, hl), lnp) in enumerate(zip(steps, lnps)):
        if not np.isfinite(lnp):
            break
        if lnp < best - 1.:
            break
        if lnp > best:
            best = lnp
            ibest = i
    if ibest is None:
        return None, 0., 0., False
    return ibest, best - lnp0, steps[ibest][0], steps[ibest][3]


def loop_should_stop(step_limited, dlnp, dchisq=0., dchisq_limited=1e-6):
    """The stopping rule of ConstrainedOptimizer.optimize_loop
    (constrained_optimizer.py:28-31): a step-limited iteration (a max-step
    clamp, or no step to try at all, line 52-55) stops only when it gained at
    most dchisq_limited; any other stops when it gained at most dchisq."""
    if step_limited:
        return dlnp <= dchisq_limited
    return dlnp <= dchisq


# ---------------------------------------------------------------------------
# Per-source fit problem (the batched configurations): one source on its own
# modelMask patch, PixelizedPSF, FFT path for galaxies, parameters
# (x, y, flux[, logre, ee1, ee2]) with an EllipseESoft shape.
# ---------------------------------------------------------------------------
class SourceProblem(object):
    """A single-source fit on a P x P patch: modelMask = (0, 0, P, P),
    positions in patch pixels, NullPhotoCal counts = flux.

    kind: 0 point source, 1 ExpGalaxy, 2 DevGalaxy, 3 SersicGalaxy (params
          (x, y, flux, logre, ee1, ee2, sersic index)).
    cdi:  2x2 position-derivative matrix (ducks.py:413-423): NullWCS gives
          3600/pixscale * I (the reference's NullWCS factor), a pixel-unit
          WCS gives I.
    hybrid_mog: (amp[K], mean[K, 2], var[K, 2, 2]) of a HybridPixelizedPSF's
          mixture (psf.py:537-625): galaxies then split their components
          between the FFT and pixel space (galaxy.py:412-455, 510-516); the
          modelMask is rectangular, so the pixel part is the FP64 grid.
    """
    def __init__(self, kind, params, P, psf_img, data=None, inverr=None, cdi=None,
                 hybrid_mog=None):
        self.kind = kind
        self.params = np.array(params, np.float64)
        self.P = P
        self.psf = np.asarray(psf_img, np.float32)
        self.data = data
        self.inverr = inverr
        self.cdi = np.eye(2) if cdi is None else np.asarray(cdi, np.float64)
        self.hybrid_mog = hybrid_mog

    @property
    def nparams(self):
        return 3 if self.kind == 0 else (7 if self.kind == 3 else 6)

    def maxstep(self):
        # PixPos maxstep 1 px (wcs.py:384); Flux none; EllipseESoft inherits
        # EllipseE maxstep [1, 0.1, 0.1] (ellipses.py:48, 190-197);
        # SersicIndex 0.1 (sersic.py:344-352)
        return [1., 1., None] + ([1., 0.1, 0.1] if self.kind else []) + (
            [0.1] if self.kind == 3 else [])

    def bounds(self):
        n = self.nparams
        if self.kind == 3:
            # SersicIndex bounds [0.29, 6.3] (sersic.py:348-351)
            return [None] * 6 + [0.29], [None] * 6 + [6.3]
        return [None] * n, [None] * n


def _profile(kind, sindex=None):
    if kind == 3:
        amp, var = sersic_mixture(sindex)
        return amp, _iso_var(var)
    return exp_mixture() if kind == 1 else dev_mixture()


# ---------------------------------------------------------------------------
# Sersic profiles (tractor/sersic.py:25-342, 447-487).  The tabulated fits
# are data, read from tractor_amd/sersic_table.py; the interpolation is
# restated here with scipy's CubicSpline (not-a-knot), the same interpolant
# as the reference's InterpolatedUnivariateSpline(k=3) through the nodes.
# ---------------------------------------------------------------------------
_SERSIC = None


def _sersic_splines():
    global _SERSIC
    if _SERSIC is None:
        from scipy.interpolate import CubicSpline, interp1d
        from tractor_amd.sersic_table import CORES, RANGES

        def spl(x, y):
            if len(x) > 3:
                return CubicSpline(x, y, bc_type='not-a-knot')
            return interp1d(x, y, fill_value='extrapolate')
        ranges = []
        for lo, hi, nodes in RANGES:
            x = [nd[0] for nd in nodes]
            K = len(nodes[0][1])
            amp = [spl(x, [nd[1][k] for nd in nodes]) for k in range(K)]
            lv = [spl(x, [math.log(nd[2][k]) for nd in nodes]) for k in range(K)]
            ranges.append((lo, hi, amp, lv))
        core = spl([c[0] for c in CORES], [c[1] for c in CORES])
        _SERSIC = (ranges, core)
    return _SERSIC


def sersic_mixture(n):
    """SersicMixture._getProfile (sersic.py:286-342) -> (amp, var): indices
    clamped to the table, amplitudes normalised per range (ramped where two
    ranges overlap), variances exp(log-variance), a zero-variance core above
    n = 1, sorted by variance."""
    ranges, core = _sersic_splines()
    if n <= ranges[0][0]:
        found, n = [ranges[0]], ranges[0][0]
    elif n >= ranges[-1][1]:
        found, n = [ranges[-1]], ranges[-1][1]
    else:
        found = [r for r in ranges if r[0] <= n < r[1]]
    if len(found) == 2:
        (lo0, hi0, a0, v0), (lo1, hi1, a1, v1) = found
        f = (n - lo1) / (hi0 - lo1)
        x0 = np.array([float(a(n)) for a in a0])
        x1 = np.array([float(a(n)) for a in a1])
        amp = np.append((1. - f) * x0 / x0.sum(), f * x1 / x1.sum())
        var = np.exp([float(v(n)) for v in v0 + v1])
    else:
        _, _, a, v = found[0]
        amp = np.array([float(f(n)) for f in a])
        amp = amp / amp.sum()
        var = np.exp([float(f(n)) for f in v])
    if n > 1.:
        c = float(core(n))
        amp = np.append(amp * (1. - c) / amp.sum(), c)
        var = np.append(var, 0.)
    o = np.argsort(var, kind='stable')
    return amp[o], var[o]


def sersic_fd_step(n0, step=0.01, lower=0.29, upper=6.3):
    """Signed finite-difference step of the Sersic index for its Jacobian
    column, as SersicGalaxy.getDerivativeShearedProfiles (sersic.py:447-487)
    takes it: +step, or -step past the upper bound; where the stepped profile
    changes the number of components, the opposite step if it stays inside
    the bounds and keeps the count."""
    if n0 + step > upper:
        step = -step
    K0 = len(sersic_mixture(n0)[0])
    if len(sersic_mixture(n0 + step)[0]) != K0:
        n2 = n0 - step
        if lower < n2 < upper and len(sersic_mixture(n2)[0]) == K0:
            step = -step
    return step


def source_render(prob, params, shape_step=None, index_step=0.):
    """Unit-flux render of the problem's source at ``params`` on its P x P
    modelMask (galaxy.py:336-508 / psf.py:137-207); a Sersic galaxy's index
    offset by ``index_step``."""
    P = prob.P
    mm = (0, 0, P, P)
    x, y = params[0], params[1]
    if prob.kind == 0:
        r = pixpsf_point_patch(prob.psf, x, y, mm=mm, derivs=False)
        return r[2] if r is not None else np.zeros((P, P), np.float32)
    sp = list(params[3:6])
    if shape_step is not None:
        sp[shape_step] += 0.01
    amp, var = _profile(prob.kind, params[6] + index_step if prob.kind == 3 else None)
    basis = ellipse_esoft_basis(*sp)
    cdinv = np.eye(2) * 3600.  # NullWCS pixscale 1 (wcs.py:24-28)
    vp = shear_mixture(var, cdinv, basis)
    r = render_galaxy_fft(amp, vp, x, y, prob.psf, mm=mm, force_sourceOut=True,
                          hybrid_mog=prob.hybrid_mog)
    return r[2]


def source_jacobian(prob, params):
    """(patch0, derivative columns as P x P float32 arrays)."""
    counts = params[2]
    if prob.kind == 0:
        mm = (0, 0, prob.P, prob.P)
        r = pixpsf_point_patch(prob.psf, params[0], params[1], mm=mm, derivs=True)
        p0, pdx, pdy = r[2], r[3], r[4]
        return p0, pointsource_param_derivs(p0, pdx, pdy, counts, prob.cdi)
    p0, d = galaxy_param_derivs(
        lambda k: source_render(prob, params, shape_step=k), counts, prob.cdi,
        [0.01, 0.01, 0.01])
    if prob.kind == 3:
        step = sersic_fd_step(params[6])
        pk = source_render(prob, params, index_step=step)
        d.append(((pk - p0) * (counts / step)).astype(np.float32))
    return p0, d


def source_chisq(prob, params, patch=None):
    """chi = (data - counts*unit) * inverr in float32; chi^2 summed in FP64
    (engine.py:455-479)."""
    if patch is None:
        patch = source_render(prob, params)
    mod = (patch * np.float32(params[2])).astype(np.float32)
    chi = ((prob.data - mod) * prob.inverr).astype(np.float32)
    return float((chi.astype(np.float64) ** 2).sum())


def source_update(prob, params):
    """Gauss-Newton direction X (column scaled, FP64 lstsq) for the current
    params (lsqr_optimizer.py:328-644 / gpu_optimizer.py:467-811)."""
    p0, derivs = source_jacobian(prob, params)
    A = np.stack([(d * prob.inverr).astype(np.float32).ravel() for d in derivs], axis=1)
    mod = (p0 * np.float32(params[2])).astype(np.float32)
    b = ((prob.data - mod) * prob.inverr).astype(np.float32).ravel()
    return column_scaled_lstsq(A, b), p0


def fit_source(prob, params0, dchisq=0., steps=100, dchisq_limited=1e-6, alphas=None):
    """ConstrainedOptimizer.optimize_loop (constrained_optimizer.py:12-34)
    with tryUpdates / getParameterSteps (36-159) and the dense column-scaled
    update.  Returns (params, n_iterations, chisq)."""
    params = np.array(params0, np.float64)
    lowers, uppers = prob.bounds()
    it = 0
    for it in range(steps):
        X, p0patch = source_update(prob, params)
        if X is None:
            break
        lnp0 = -0.5 * source_chisq(prob, params, p0patch)
        flags = {'step_limited': False}
        stps = parameter_steps(params, X, prob.maxstep(), lowers, uppers, alphas, flags=flags)
        # tryUpdates marks an iteration with nothing to try as step-limited
        # (constrained_optimizer.py:52-55)
        step_limited = flags['step_limited'] or len(stps) == 0
        lnps = []
        # the reference evaluates sequentially and stops early; evaluating
        # every step and applying the same rule gives the same choice.
        for (alpha, p, sl, hl) in stps:
            lnps.append(-0.5 * source_chisq(prob, np.array(p)))
        ib, dlnp, alpha, hl = accept_line_search(lnp0, stps, lnps)
        if ib is not None:
            params = np.array(stps[ib][1], np.float64)
        if loop_should_stop(step_limited, dlnp, dchisq, dchisq_limited):
            break
    return params, it + 1, source_chisq(prob, params)


# ---------------------------------------------------------------------------
# One galaxy in a whole image without modelMasks (the single-galaxy scene):
# LsqrOptimizer.optimize_loop (lsqr_optimizer.py:280-326) with the
# unconstrained line search of Optimizer.tryUpdates (optimize.py:425-470).
# ---------------------------------------------------------------------------
def image_galaxy_unit_patch(kind, params, psf_img, img_shape, mm=None):
    """Unit-flux patch (x0, y0, float32 patch) of an Exp (kind 1) / Dev
    (kind 2) galaxy with EllipseESoft shape params[3:6] at pixel (params[0],
    params[1]), NullWCS(pixscale=1) shapes, PixelizedPSF psf_img.  Without a
    window the extent is the reference's no-modelMask rule: halfsize = nre r_e
    (nre 4 exp, 8 dev; galaxy.py:728-756), at most the distance to the
    farthest image edge, at least the PSF radius hypot(H/2, W/2) (psf.py:97),
    rounded up (galaxy.py:678-691); pixels [floor(p - halfsize),
    ceil(p + halfsize + 1)] clipped to the image (galaxy.py:273-290).  With
    mm = (x0, y0, w, h): rendered on that window with force_sourceOut, as the
    shape-stepped renders of getParamDerivatives are (galaxy.py:199-205).
    None when the patch misses the image."""
    amp, var = _profile(kind)
    vp = shear_mixture(var, np.eye(2) * 3600., ellipse_esoft_basis(*params[3:6]))
    px, py = float(params[0]), float(params[1])
    if mm is not None:
        return render_galaxy_fft(amp, vp, px, py, psf_img, mm=mm, force_sourceOut=True)
    ph, pw = np.shape(psf_img)
    nre = 4. if kind == 1 else 8.
    hs = galaxy_patch_halfsize(nre * math.exp(params[3]), px, py, img_shape,
                               math.hypot(ph / 2., pw / 2.))
    H, W = img_shape
    ox, ix = get_overlapping_region(int(math.floor(px - hs)), int(math.ceil(px + hs + 1)), 0, W - 1)
    oy, iy = get_overlapping_region(int(math.floor(py - hs)), int(math.ceil(py + hs + 1)), 0, H - 1)
    if ix == [] or iy == []:
        return None
    return render_galaxy_fft(amp, vp, px, py, psf_img,
                             nomask_extent=(ox.start, ox.stop, oy.start, oy.stop, hs))


def image_model(kind, params, psf_img, img_shape):
    """float32 model image: counts x unit patch added into zeros
    (engine.py:424-445; counts = flux, NullPhotoCal)."""
    mod = np.zeros(img_shape, np.float32)
    r = image_galaxy_unit_patch(kind, params, psf_img, img_shape)
    if r is not None:
        x0, y0, p = r
        mod[y0:y0 + p.shape[0], x0:x0 + p.shape[1]] += (p * np.float32(params[2]))
    return mod


def image_chisq(kind, params, psf_img, data, inverr):
    """chi^2 of the whole image: float32 chi, FP64 sum of squares
    (engine.py:455-479)."""
    chi = ((data - image_model(kind, params, psf_img, data.shape)) * inverr).astype(np.float32)
    return float((chi.astype(np.float64) ** 2).sum())


def fit_image_lsqr(kind, params0, psf_img, data, inverr, cdi, dchisq=1e-3, steps=50,
                   alphas=None):
    """LsqrOptimizer.optimize_loop on one galaxy of a whole image, all six
    parameters thawed, no priors: per step the column-scaled least-squares
    direction of the derivative columns (galaxy.py:137-223 with cdi the
    position-derivative matrix, ducks.py:413-423) over the patch pixels, then
    every alpha of Optimizer.tryUpdates walked by accept_line_search; the
    loop stops after a step that gains at most dchisq (lsqr_optimizer.py:
    317-326).  Returns (params, last step index, chisq, [(dlnp, alpha), ...])."""
    alphas = ALPHAS if alphas is None else np.asarray(alphas)
    params = np.array(params0, np.float64)
    trace = []
    step = 0
    for step in range(steps):
        x0, y0, p0 = image_galaxy_unit_patch(kind, params, psf_img, data.shape)
        h, w = p0.shape

        def render(k):
            if k is None:
                return p0
            q = params.copy()
            q[3 + k] += 0.01
            return image_galaxy_unit_patch(kind, q, psf_img, data.shape, mm=(x0, y0, w, h))[2]
        _, derivs = galaxy_param_derivs(render, params[2], cdi, [0.01, 0.01, 0.01])
        ie = inverr[y0:y0 + h, x0:x0 + w]
        A = np.stack([(d * ie).astype(np.float32).ravel() for d in derivs], axis=1)
        chi = ((data - image_model(kind, params, psf_img, data.shape)) * inverr).astype(np.float32)
        X = column_scaled_lstsq(A, chi[y0:y0 + h, x0:x0 + w].ravel())
        if X is None:
            break
        lnp0 = -0.5 * float((chi.astype(np.float64) ** 2).sum())
        trials = [(a, params + a * X, False, False) for a in alphas]
        lnps = [-0.5 * image_chisq(kind, p, psf_img, data, inverr) for (_, p, _, _) in trials]
        ib, dlnp, alpha, _ = accept_line_search(lnp0, trials, lnps)
        if ib is not None:
            params = trials[ib][1]
        trace.append((dlnp, alpha))
        if dlnp <= dchisq:
            break
    return params, step, image_chisq(kind, params, psf_img, data, inverr), trace


# ---------------------------------------------------------------------------
# Forced photometry (optimize.py:16-222 with LsqrOptimizer's unscaled step,
# lsqr_optimizer.py:11-251) of many sources with per-source pixelized PSFs.
# ---------------------------------------------------------------------------
def forced_unit_patch(kind, px, py, shape, stamp, img_shape):
    """Unit-flux patch of one source without a modelMask, clipped to the
    image: point sources via PixelizedPSF.getPointSourcePatch (psf.py:137-207,
    the Lanczos-shifted stamp), exp (kind 1) / dev (kind 2) galaxies with
    EllipseESoft shape (log r_e, ee1, ee2) via _getUnitFluxPatchSize
    (galaxy.py:678-691) and the FFT path (galaxy.py:336-508).  Returns
    (x0, y0, patch float64) or None."""
    H, W = img_shape
    stamp = np.asarray(stamp, np.float32)
    ph, pw = stamp.shape
    if kind == 0:
        x0, y0, p = pixpsf_point_patch(stamp, px, py)
    else:
        amp, var = exp_mixture() if kind == 1 else dev_mixture()
        var = shear_mixture(var, np.eye(2) * 3600., ellipse_esoft_basis(*shape))
        nre = 4. if kind == 1 else 8.
        hs = galaxy_patch_halfsize(nre * math.exp(shape[0]), px, py, img_shape,
                                   math.hypot(ph / 2., pw / 2.))
        outx, _ = get_overlapping_region(int(np.floor(px - hs)), int(np.ceil(px + hs + 1)), 0, W - 1)
        outy, _ = get_overlapping_region(int(np.floor(py - hs)), int(np.ceil(py + hs + 1)), 0, H - 1)
        if outx == [] or outy == []:
            return None
        x0, y0, p = render_galaxy_fft(amp, var, px, py, stamp,
                                      nomask_extent=(outx.start, outx.stop, outy.start,
                                                     outy.stop, hs))
    p = np.asarray(p, np.float64)
    h, w = p.shape
    X0, Y0 = max(x0, 0), max(y0, 0)
    X1, Y1 = min(x0 + w, W), min(y0 + h, H)
    if X1 <= X0 or Y1 <= Y0:
        return None
    return X0, Y0, p[Y0 - y0:Y1 - y0, X0 - x0:X1 - x0]


def forced_phot_lsq(patches, data, inverr):
    """Least-squares fluxes min || (data - sum_s f_s u_s) inverr || over the
    unit-flux patches (x0, y0, patch) (None: flux 0), solved exactly in FP64
    through the sparse normal equations; returns (flux, iv) with iv =
    sum (u inverr)^2 (Optimizer._get_iv, optimize.py:378-423)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    H, W = data.shape
    rows, cols, vals = [], [], []
    for s, pt in enumerate(patches):
        if pt is None:
            continue
        x0, y0, p = pt
        h, w = p.shape
        yy, xx = np.mgrid[y0:y0 + h, x0:x0 + w]
        rows.append((yy * W + xx).ravel())
        cols.append(np.full(h * w, s))
        vals.append((p * inverr[y0:y0 + h, x0:x0 + w]).ravel())
    n = len(patches)
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(H * W, n))
    b = (np.asarray(data, np.float64) * inverr).ravel()
    N = (A.T @ A).tocsc()
    iv = N.diagonal().copy()
    empty = iv == 0
    N = N + sp.diags(empty.astype(np.float64))
    f = spl.spsolve(N, A.T @ b)
    f[empty] = 0.
    return f, iv


def forced_fitstats(patches, counts, model, data, inverr, scale=1., sky=0.):
    """Optimizer._get_fitstats (optimize.py:233-376) for one image and one
    unit-flux model per source, FP64.  patches: (x0, y0, u) or None per
    source; counts: each source's counts (flux x photocal scale); model: the
    image model without sky (sum_s counts_s u_s plus fixed sources); chi =
    (data - (model + sky)) inverr.  Returns per-source arrays prochi2,
    pronpix, promasked, profracflux, proflux, npix, fracin and the image
    imchisq, imnpix; also iv = sum (u scale inverr)^2 (_get_iv,
    optimize.py:378-423).  Over several images the sums add, except
    profracflux = sum_img(num / den2) / max(1, sum_img den) and fracin =
    sum_img fracin / max(1, number of images counted): the per-image parts
    are returned as well (ffnum, ffden, counted)."""
    model = np.asarray(model, np.float64)
    data = np.asarray(data, np.float64)
    inverr = np.asarray(inverr, np.float64)
    chi = (data - (model + sky)) * inverr
    n = len(patches)
    z = lambda: np.zeros(n)
    fs = dict(prochi2=z(), pronpix=z(), promasked=z(), ffnum=z(), ffden=z(), proflux=z(),
              npix=np.zeros(n, int), fracin_num=z(), counted=z(), iv=z())
    for s, pt in enumerate(patches):
        if pt is None:
            continue
        x0, y0, u = pt
        u = np.asarray(u, np.float64)
        h, w = u.shape
        sl = slice(y0, y0 + h), slice(x0, x0 + w)
        ie = inverr[sl]
        fs['iv'][s] = np.sum((u * scale * ie) ** 2)
        c = counts[s]
        if c == 0:
            continue
        nz = (u != 0) & (ie > 0)
        if not nz.any():
            continue
        au = np.abs(u)
        m = model[sl]
        fs['prochi2'][s] = np.sum(au[nz] * chi[sl][nz] ** 2)
        fs['pronpix'][s] = np.sum(au[nz])
        fs['promasked'][s] = np.sum(au[ie == 0])
        fs['ffnum'][s] = np.sum((np.abs(m / c - u) * au)[nz]) / np.sum((u ** 2)[nz])
        fs['ffden'][s] = np.sum(au[nz]) / abs(c)
        fs['proflux'][s] = np.sum((np.abs((m - u * c) / scale) * au)[nz])
        fs['npix'][s] = int(nz.sum())
        fs['fracin_num'][s] = np.sum(au)
        fs['counted'][s] = 1.
    fs['profracflux'] = fs['ffnum'] / np.maximum(1., fs['ffden'])
    fs['fracin'] = fs['fracin_num'] / np.maximum(1., fs['counted'])
    fs['imchisq'] = float(np.sum(chi ** 2))
    fs['imnpix'] = int(np.sum(inverr > 0))
    return fs


def psfex_terms(x, y, degree, x0, y0, xscale, yscale):
    """PsfEx.polynomials (psfex.py:294-325): the term dx^j dy^k, dx = (x - x0)
    / xscale, dy = (y - y0) / yscale, sits at index
    j + (degree + 1) k - k (k - 1) / 2 for every j + k <= degree."""
    dx = (x - x0) / xscale
    dy = (y - y0) / yscale
    nb = (degree + 1) * (degree + 2) // 2
    terms = np.zeros(nb)
    for d in range(degree + 1):
        for j in range(d + 1):
            k = d - j
            terms[j + (degree + 1) * k - (k * (k - 1)) // 2] = dx ** j * dy ** k
    return terms


def psfex_at(bases, x, y, degree, x0, y0, xscale, yscale):
    """PsfEx.at (psfex.py:326-346): sum of the eigen-PSFs weighted by the
    polynomial terms at (x, y), FP64."""
    t = psfex_terms(x, y, degree, x0, y0, xscale, yscale)
    return np.tensordot(t, np.asarray(bases, np.float64), axes=(0, 0))


def em_fit_2d_reg(img, x0, y0, amp, mean, var, alpha, steps):
    """em_fit_2d_reg (emfit.i:274-...), FP64: EM fit of a K-component 2-D
    mixture of Gaussians to image weights, with an inverse-Wishart prior of
    strength alpha on the variances (prior I).  Pixel (ix, iy) sits at
    (x0 + ix, y0 + iy).  E step: z_k = amp_k N(x | mean_k, var_k), terms with
    Mahalanobis distance^2 >= 100 set to 0, normalised over k (pixels with
    all terms 0 are dropped); M step: wsum_k = sum img z_k, mean_k = the
    img z_k weighted centroid, var_k = (alpha I + sum img z_k d d^T) /
    (wsum_k + alpha) about the new mean, amp_k = wsum_k.  Returns (amp, mean,
    var) (copies); raises on a non-positive determinant."""
    img = np.asarray(img, np.float64)
    amp = np.array(amp, np.float64)
    mean = np.array(mean, np.float64).reshape(-1, 2)
    var = np.array(var, np.float64).reshape(-1, 2, 2)
    K = len(amp)
    NY, NX = img.shape
    yy, xx = np.mgrid[0:NY, 0:NX]
    X = (x0 + xx).ravel().astype(np.float64)
    Y = (y0 + yy).ravel().astype(np.float64)
    w = img.ravel()
    tpd = (2. * np.pi) ** 2
    for _ in range(steps):
        Z = np.zeros((K, X.size))
        for k in range(K):
            V = var[k]
            det = V[0, 0] * V[1, 1] - V[0, 1] * V[1, 0]
            if det <= 0.:
                raise ValueError('non-positive determinant')
            iv = np.array([V[1, 1], -V[0, 1], -V[1, 0], V[0, 0]]) / det
            dx, dy = X - mean[k, 0], Y - mean[k, 1]
            dsq = iv[0] * dx * dx + iv[1] * dx * dy + iv[2] * dx * dy + iv[3] * dy * dy
            z = amp[k] / np.sqrt(tpd * det) * np.exp(-0.5 * dsq)
            Z[k] = np.where(dsq >= 100., 0., z)
        zs = Z.sum(0)
        ok = zs != 0
        Z[:, ok] /= zs[ok]
        wz = w[None, :] * Z
        wsum = wz.sum(1)
        mean = np.stack([(wz * X).sum(1), (wz * Y).sum(1)], 1) / wsum[:, None]
        for k in range(K):
            dx, dy = X - mean[k, 0], Y - mean[k, 1]
            S = np.array([[alpha + np.sum(wz[k] * dx * dx), np.sum(wz[k] * dx * dy)],
                          [np.sum(wz[k] * dy * dx), alpha + np.sum(wz[k] * dy * dy)]])
            var[k] = S / (wsum[k] + alpha)
        amp = wsum.copy()
    return amp, mean, var


def em_init_params(K):
    """fitpsf.em_init_params(K, None, None, None) (fitpsf.py:72-86): equal
    weights, zero means, variances I (k + 1)."""
    return (np.ones(K) / K, np.zeros((K, 2)),
            np.array([np.eye(2) * (k + 1) for k in range(K)]))
