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
