"""Synthetic workloads of the batched render-and-fit benchmarks.

Seeds follow numpy.random.default_rng(20261203 + 1000*config + k), with
separate streams for source parameters (k=0), start perturbations (k=1) and
noise (k=2, drawn on the device by torch's generator with that seed).
"""
import math

import numpy as np

KIND_POINT, KIND_EXP, KIND_DEV = 0, 1, 2


def seed(config, k):
    return 20261203 + 1000 * config + k


def moffat_stamp(size=63, fwhm=4.0, beta=3.5):
    """Pixel-centre samples of a circular Moffat profile centred on the middle
    pixel, normalised to unit sum."""
    alpha = fwhm / (2. * math.sqrt(2. ** (1. / beta) - 1.))
    c = size // 2
    yy, xx = np.mgrid[:size, :size].astype(np.float64)
    r2 = (xx - c) ** 2 + (yy - c) ** 2
    img = (1. + r2 / alpha ** 2) ** (-beta)
    return (img / img.sum()).astype(np.float32)


def gaussian_mixture_stamp(size, sigmas, weights):
    """Pixel-centre samples of concentric circular Gaussians, unit sum."""
    c = size // 2
    yy, xx = np.mgrid[:size, :size].astype(np.float64)
    r2 = (xx - c) ** 2 + (yy - c) ** 2
    img = sum(w * np.exp(-0.5 * r2 / s ** 2) / (2 * np.pi * s ** 2) for s, w in zip(sigmas, weights))
    return (img / img.sum()).astype(np.float32)


def config1_scene(size=64, truth_e=(0.2, -0.1)):
    """The single-galaxy scene: an ExpGalaxy near the centre of a size x size
    image, flux 1000, r_e 3 px, EllipseE (e1, e2) = truth_e carried as
    EllipseESoft (log r_e, ee1, ee2); 31x31 pixel-centre stamp of
    0.9 N(sigma 1.5) + 0.1 N(sigma 4.0), unit sum; unit-variance noise.
    Start: position + N(0, 0.5) px, flux and r_e x U(0.8, 1.2).
    Returns dict(psf, truth, start, noise), parameters (x, y, flux, log r_e,
    ee1, ee2)."""
    rng = np.random.default_rng(seed(1, 0))
    u = rng.uniform(0., 1.)
    re = 3.
    e1, e2 = truth_e
    e = math.hypot(e1, e2)
    # EllipseESoft.fromEllipseE (ellipses.py:199-203): |ee| = -log(1 - |e|),
    # same direction
    ee = -math.log(1. - e) / e if e > 0 else 1.
    truth = np.array([31.5 + u, 31.5 + u, 1000., math.log(re), e1 * ee, e2 * ee])
    r1 = np.random.default_rng(seed(1, 1))
    start = truth.copy()
    start[:2] += r1.normal(0., 0.5, 2)
    start[2] *= r1.uniform(0.8, 1.2)
    start[3] = math.log(re * r1.uniform(0.8, 1.2))
    noise = np.random.default_rng(seed(1, 2)).normal(0., 1., (size, size)).astype(np.float32)
    return dict(psf=gaussian_mixture_stamp(31, (1.5, 4.0), (0.9, 0.1)), truth=truth,
                start=start, noise=noise)


def draw_sources(n, config=2, P_choices=None, frac=(0.3, 0.4, 0.3), rng=None):
    """Source catalogue of the batched configurations: kind, patch size P,
    truth params (x, y, flux, logre, ee1, ee2) in patch pixels.

    30% PointSource / 40% ExpGalaxy / 30% DevGalaxy; P uniform on
    {16, 24, ..., 128}; flux log-uniform 50..1e5; r_e log-uniform 0.5..10 px;
    |e| ~ U(0, 0.6), position angle ~ U(0, pi); centre P//2 + U[0,1)^2.
    """
    if rng is None:
        rng = np.random.default_rng(seed(config, 0))
    if P_choices is None:
        P_choices = np.arange(16, 129, 8)
    npt = int(round(frac[0] * n))
    nexp = int(round(frac[1] * n))
    kind = np.array([KIND_POINT] * npt + [KIND_EXP] * nexp + [KIND_DEV] * (n - npt - nexp),
                    np.int32)
    rng.shuffle(kind)
    P = rng.choice(P_choices, size=n).astype(np.int32)
    params = np.zeros((n, 6))
    params[:, 0] = P // 2 + rng.uniform(0, 1, n)
    params[:, 1] = P // 2 + rng.uniform(0, 1, n)
    params[:, 2] = np.exp(rng.uniform(np.log(50.), np.log(1e5), n))
    re = np.exp(rng.uniform(np.log(0.5), np.log(10.), n))
    e = rng.uniform(0., 0.6, n)
    th = rng.uniform(0., np.pi, n)
    e1, e2 = e * np.cos(2 * th), e * np.sin(2 * th)
    for i in range(n):
        if kind[i] == KIND_POINT:
            continue
        ee = -math.log(1. - e[i])
        if e[i] > 0:
            params[i, 3:6] = (math.log(re[i]), e1[i] / e[i] * ee, e2[i] / e[i] * ee)
        else:
            params[i, 3:6] = (math.log(re[i]), 0., 0.)
    return kind, P, params


def perturb(kind, params, rng, pos=0.5, frac=0.3, de=0.1):
    """Start point: truth + U(-pos, pos) px, flux and r_e x U(1-frac, 1+frac),
    e1/e2 + U(-de, de) (applied to the unsoftened ellipticity)."""
    n = len(kind)
    p = params.copy()
    p[:, 0] += rng.uniform(-pos, pos, n)
    p[:, 1] += rng.uniform(-pos, pos, n)
    p[:, 2] *= rng.uniform(1 - frac, 1 + frac, n)
    r = rng.uniform(1 - frac, 1 + frac, n)
    d1 = rng.uniform(-de, de, n)
    d2 = rng.uniform(-de, de, n)
    for i in range(n):
        if kind[i] == KIND_POINT:
            p[i, 3:6] = 0.
            continue
        logre, ee1, ee2 = params[i, 3:6]
        ee = math.hypot(ee1, ee2)
        e = 1. - math.exp(-ee)
        e1, e2 = (e * ee1 / ee, e * ee2 / ee) if ee > 0 else (0., 0.)
        e1, e2 = e1 + d1[i], e2 + d2[i]
        e = min(math.hypot(e1, e2), 0.95)
        th = math.atan2(e2, e1)
        eesoft = -math.log(1. - e)
        p[i, 3:6] = (logre + math.log(r[i]), eesoft * math.cos(th), eesoft * math.sin(th))
    return p


def patch_offsets(P):
    off = np.zeros(len(P), np.int64)
    off[1:] = np.cumsum(P.astype(np.int64) ** 2)[:-1]
    return off, int(np.sum(P.astype(np.int64) ** 2))


def psf_grid_stamps(size=33, fwhm_lo=3.0, fwhm_hi=5.0, ngrid=3):
    """ngrid x ngrid grid of PSF stamps for a spatially varying PSF: two
    concentric Gaussians (sigma, 2 sigma; weights 0.85, 0.15), FWHM rising
    linearly from fwhm_lo at node (0, 0) to fwhm_hi at the far corner.
    Returns (ngrid*ngrid, size, size) float32, node (gx, gy) at gy*ngrid+gx."""
    out = []
    for gy in range(ngrid):
        for gx in range(ngrid):
            f = fwhm_lo + (fwhm_hi - fwhm_lo) * (gx + gy) / (2. * (ngrid - 1))
            s = f / (2. * math.sqrt(2. * math.log(2.)))
            out.append(gaussian_mixture_stamp(size, [s, 2. * s], [0.85, 0.15]))
    return np.array(out, np.float32)


def bilinear_psf_weights(x, y, W, H, ngrid=3):
    """Bilinear interpolation weights of the PSF grid (nodes at
    i (W-1)/(ngrid-1), j (H-1)/(ngrid-1)) at pixel positions (x, y):
    returns (idx (n, 4) int32, w (n, 4) float32)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    u = x / ((W - 1.) / (ngrid - 1))
    v = y / ((H - 1.) / (ngrid - 1))
    i0 = np.clip(np.floor(u).astype(int), 0, ngrid - 2)
    j0 = np.clip(np.floor(v).astype(int), 0, ngrid - 2)
    tx = np.clip(u - i0, 0., 1.)
    ty = np.clip(v - j0, 0., 1.)
    idx = np.stack([j0 * ngrid + i0, j0 * ngrid + i0 + 1, (j0 + 1) * ngrid + i0,
                    (j0 + 1) * ngrid + i0 + 1], 1).astype(np.int32)
    w = np.stack([(1 - tx) * (1 - ty), tx * (1 - ty), (1 - tx) * ty, tx * ty], 1).astype(np.float32)
    return idx, w


def draw_forced_scene(n=20000, size=2048, frac_point=0.6, rng=None):
    """Forced-photometry scene: n sources at Poisson (uniform) positions on a
    size x size image; 60% point sources, the rest exp / dev galaxies (4:3)
    with fixed shapes drawn as the batched benchmark (r_e log-uniform
    0.5..10 px, |e| ~ U(0, 0.6), angle ~ U(0, pi)); fluxes log-uniform
    10..1e4.  Returns kind (n,), pos (n, 2), flux (n,), shape (n, 3) =
    EllipseESoft (log r_e, ee1, ee2)."""
    if rng is None:
        rng = np.random.default_rng(seed(3, 0))
    npt = int(round(frac_point * n))
    ngal = n - npt
    nexp = int(round(ngal * 4. / 7.))
    kind = np.array([KIND_POINT] * npt + [KIND_EXP] * nexp + [KIND_DEV] * (ngal - nexp), np.int32)
    rng.shuffle(kind)
    pos = rng.uniform(0., size, (n, 2))
    flux = np.exp(rng.uniform(np.log(10.), np.log(1e4), n))
    re = np.exp(rng.uniform(np.log(0.5), np.log(10.), n))
    e = rng.uniform(0., 0.6, n)
    th = rng.uniform(0., np.pi, n)
    ee = -np.log(1. - e)
    shape = np.stack([np.log(re), ee * np.cos(2 * th), ee * np.sin(2 * th)], 1)
    shape[kind == KIND_POINT] = 0.
    return kind, pos, flux, shape
