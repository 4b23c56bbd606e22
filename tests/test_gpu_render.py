"""GPU parity: libtractor_hip render kernels vs the FP64 oracle.

Tolerance (the parity bar for FP32 renders): max |gpu - oracle| /
max |oracle| <= 1e-5 per patch.
"""
import math
import sys

import numpy as np
import pytest
import torch

import oracle as O
from tractor_amd import render as R
from tractor_amd import workloads as W
from tractor_amd._lib import FFT_JOB, TRH_STORE, TRH_CHI2

pytestmark = pytest.mark.gpu
RTOL = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope='module')
def dev():
    assert torch.cuda.is_available(), 'GPU tests need a HIP device'
    return 'cuda'


def rel_err(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


@pytest.mark.parametrize('sz', [32, 64, 128, 256, 512])
def test_psf_spectrum(dev, sz):
    say('psf spectrum sz', sz)
    stamp = W.moffat_stamp(63, 4.0, 3.5)
    sp = R.PsfSpectra(stamp, dev)
    spec, cx, cy = sp.get(sz)
    torch.cuda.synchronize()
    got = spec.cpu().numpy()
    got = (got[..., 0] + 1j * got[..., 1]).T     # stored [sz/2+1][sz]
    P, (ocx, ocy), _, _ = O.psf_spectrum(stamp, sz / 2.)
    assert (cx, cy) == (ocx, ocy)
    assert rel_err(got, P) < RTOL


def _random_galaxies(n, rng, sizes=(16, 24, 40, 56, 64, 88, 120, 128)):
    out = []
    for i in range(n):
        kind = 1 + (i % 2)
        P = int(sizes[i % len(sizes)])
        x = P // 2 + rng.uniform(0, 1)
        y = P // 2 + rng.uniform(0, 1)
        re = math.exp(rng.uniform(math.log(0.5), math.log(10.)))
        e = rng.uniform(0, 0.6)
        th = rng.uniform(0, math.pi)
        out.append((kind, P, x, y, O.esoft_from_e(re, e * math.cos(2 * th), e * math.sin(2 * th))))
    return out


def _sheared(kind, shape):
    amp, var = O.exp_mixture() if kind == 1 else O.dev_mixture()
    return amp, O.shear_mixture(var, np.eye(2) * 3600., O.ellipse_esoft_basis(*shape))


def test_fft_render_vs_oracle(dev):
    rng = np.random.default_rng(7)
    psf = W.moffat_stamp(63, 4.0, 3.5)
    sp = R.PsfSpectra(psf, dev)
    gals = _random_galaxies(24, rng)
    by_sz = {}
    offs = []
    total = 0
    for (kind, P, x, y, shape) in gals:
        amp, var = _sheared(kind, shape)
        sz, J, win = R.galaxy_fft_job(amp, var, x, y, psf.shape, mm=(0, 0, P, P))
        J['out_offset'] = total
        J['out_stride'] = P
        J['mode'] = TRH_STORE
        by_sz.setdefault(sz, []).append(J)
        offs.append(total)
        total += P * P
    out = torch.zeros(total, dtype=torch.float32, device=dev)
    for sz, js in by_sz.items():
        say('fft render sz', sz, 'jobs', len(js))
        R.render_fft_jobs(sz, np.array(js, dtype=FFT_JOB), sp, out=out, device=dev)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    worst = 0.
    for (kind, P, x, y, shape), off in zip(gals, offs):
        amp, var = _sheared(kind, shape)
        _, _, G = O.render_galaxy_fft(amp, var, x, y, psf, mm=(0, 0, P, P))
        e = rel_err(got[off:off + P * P].reshape(P, P), G)
        worst = max(worst, e)
        assert e < RTOL, (kind, P, x, y, shape, e)
    say('worst rel err', worst)


def test_fft_render_nomask_and_shifts(dev):
    # no-modelMask path: the window is the clipped extent (galaxy.py:273-290, 495-508)
    rng = np.random.default_rng(11)
    psf = W.gaussian_mixture_stamp(25, [1.5, 4.0], [0.9, 0.1])
    sp = R.PsfSpectra(psf, dev)
    Hh, Ww = 100, 100
    cases = [(49.5, 50.), (5., 5.), (95.3, 3.7), (50.2, 97.9)]
    for (px, py) in cases:
        amp, var = _sheared(1, O.esoft_from_e(3., 0.2, -0.1))
        radius = np.hypot(12.5, 12.5)
        hs = O.galaxy_patch_halfsize(4 * 3., px, py, (Hh, Ww), radius)
        outx, _ = O.get_overlapping_region(int(np.floor(px - hs)), int(np.ceil(px + hs + 1)), 0, Ww - 1)
        outy, _ = O.get_overlapping_region(int(np.floor(py - hs)), int(np.ceil(py + hs + 1)), 0, Hh - 1)
        ext = (outx.start, outx.stop, outy.start, outy.stop, hs)
        ox0, oy0, G = O.render_galaxy_fft(amp, var, px, py, psf, nomask_extent=ext)
        sz, J, win = R.galaxy_fft_job(amp, var, px, py, psf.shape, nomask_extent=ext)
        assert (win[0], win[1], win[3], win[2]) == (ox0, oy0) + G.shape
        img = torch.zeros(Hh * Ww, dtype=torch.float32, device=dev)
        J['out_offset'] = win[1] * Ww + win[0]
        J['out_stride'] = Ww
        J['mode'] = TRH_STORE
        say('nomask render', px, py, 'sz', sz)
        R.render_fft_jobs(sz, np.array([J], dtype=FFT_JOB), sp, out=img, device=dev)
        torch.cuda.synchronize()
        got = img.cpu().numpy().reshape(Hh, Ww)[win[1]:win[1] + win[3], win[0]:win[0] + win[2]]
        assert rel_err(got, G) < RTOL


def test_point_render_vs_oracle(dev):
    rng = np.random.default_rng(3)
    psf = W.moffat_stamp(63, 4.0, 3.5)
    sp = R.PsfSpectra(psf, dev)
    js, cases, total = [], [], 0
    for P in (16, 40, 64, 100, 128):
        for _ in range(3):
            x = P // 2 + rng.uniform(0, 1)
            y = P // 2 + rng.uniform(0, 1)
            J = R.point_job(x, y, psf.shape, (0, 0, P, P))
            J['out_offset'] = total
            J['out_stride'] = P
            J['mode'] = TRH_STORE
            js.append(J)
            cases.append((P, x, y, total))
            total += P * P
    out = torch.zeros(total, dtype=torch.float32, device=dev)
    say('point renders', len(js))
    R.render_point_jobs(np.array(js, dtype=FFT_JOB), sp, out=out, device=dev)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    for (P, x, y, off) in cases:
        ref = O.pixpsf_point_patch(psf, x, y, mm=(0, 0, P, P))[2]
        assert rel_err(got[off:off + P * P].reshape(P, P), ref) < RTOL, (P, x, y)


def test_chi2_epilogue(dev):
    # TRH_CHI2 = sum(((data - scale*model)*inverr)^2) in FP64 (engine.py:455-479)
    rng = np.random.default_rng(5)
    psf = W.moffat_stamp(63, 4.0, 3.5)
    sp = R.PsfSpectra(psf, dev)
    kind, P, x, y, shape = 2, 64, 32.3, 32.8, O.esoft_from_e(2.5, 0.3, 0.1)
    amp, var = _sheared(kind, shape)
    _, _, G = O.render_galaxy_fft(amp, var, x, y, psf, mm=(0, 0, P, P))
    data = (G * 1000. + rng.normal(size=G.shape)).astype(np.float32)
    ie = np.full(G.shape, 0.5, np.float32)
    sz, J, _ = R.galaxy_fft_job(amp, var, x, y, psf.shape, mm=(0, 0, P, P))
    J['out_offset'], J['out_stride'], J['mode'], J['slot'], J['scale'] = 0, P, TRH_CHI2, 1, 1000.
    chi2 = torch.zeros(2, dtype=torch.float64, device=dev)
    R.render_fft_jobs(sz, np.array([J], dtype=FFT_JOB), sp, data=torch.from_numpy(data.ravel()).to(dev),
                      inverr=torch.from_numpy(ie.ravel()).to(dev), chi2=chi2, device=dev)
    torch.cuda.synchronize()
    ref = float((((data - (G * np.float32(1000.)).astype(np.float32)) * ie).astype(np.float64) ** 2).sum())
    assert abs(chi2[1].item() - ref) / ref < 1e-5


def test_autotuner_installs_shapes_and_keeps_results(dev):
    """tune_fft times the candidate shapes and installs the fastest.  The
    launch shape moves no per-pixel arithmetic: STORE renders are identical
    bit for bit under every shape, and the FP64 chi^2 sums differ only by
    summation order (1e-12 relative)."""
    from tractor_amd import autotune as A, _lib
    from tractor_amd._lib import TRH_STORE
    say('autotune')
    stamp = W.moffat_stamp(63, 4.0, 3.5)
    sp = R.PsfSpectra(stamp, dev)
    rng = np.random.default_rng(4)
    J, npix = A._jobs(128, np.array([72, 96, 120]), 256, stamp.shape, rng)
    Js = J.copy()
    Js['mode'] = TRH_STORE
    dj = R.jobs_to_device(J, dev)
    djs = R.jobs_to_device(Js, dev)
    data = torch.randn(npix, device=dev)
    inv = torch.ones_like(data)
    lib, h = _lib.load(), _lib.handle(0)
    spec, _, _ = sp.get(128)

    def render():
        c = torch.zeros(256, dtype=torch.float64, device=dev)
        out = torch.zeros(npix, device=dev)
        _lib.check(lib.trh_render_fft_batch(h.value, 128, _lib.ptr(dj), 256, _lib.ptr(spec), None,
                                            _lib.ptr(data), _lib.ptr(inv), _lib.ptr(c), None), 'r')
        _lib.check(lib.trh_render_fft_batch(h.value, 128, _lib.ptr(djs), 256, _lib.ptr(spec),
                                            _lib.ptr(out), None, None, None, None), 'r')
        torch.cuda.synchronize()
        return c.cpu().numpy(), out.cpu().numpy()
    c0, p0 = render()
    best = A.tune_fft(stamp, [72, 96, 120], njobs=2048, reps=1)
    assert set(best) == {128} and best[128][2] > 0
    for t, r in [(None, None), (256, 64), (1024, 0), (512, 120)]:
        if t is not None:
            _lib.check(lib.trh_set_fft_config(h.value, 128, t, r), 'cfg')
        c1, p1 = render()
        assert np.array_equal(p1, p0), (t, r)
        np.testing.assert_allclose(c1, c0, rtol=1e-12, atol=0)
    _lib.check(lib.trh_set_fft_config(h.value, 128, 0, 0), 'cfg')
