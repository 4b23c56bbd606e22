"""Sources whose renders need an FFT size outside 32..512 do not fail the
batch: a Jacobian that needs one stops that source (TRH_FIT_F_JAC_SIZE), a
line-search trial that needs one counts as a non-finite log-probability
(TRH_FIT_F_LS_SIZE), and every other source's fit is unchanged bit for bit.

The FFT size follows the modelMask halfsize rule (galaxy.py:337-351): for a
P x P mask the halfsize is max(P/2, 1 + x, 1 + P - x, ...), so a P = 300
galaxy at x = 255 (inside its mask) needs sz = 512 and any trial right of
x = 255 needs 1024.  The data of such a source is rendered by the oracle
(any FFT size)."""
import math
import sys

import numpy as np
import pytest
import torch

import oracle as O
from tractor_amd import batch as B
from tractor_amd import workloads as W
from tractor_amd._lib import TRH_FIT_F_JAC_SIZE, TRH_FIT_F_LS_SIZE, TRH_FIT_F_MAXITER

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _scene():
    psf = W.moffat_stamp(63, 4.0, 3.5)
    kind = np.array([1, 2, 0, 1, 2], np.int32)
    P = np.array([48, 64, 32, 300, 300], np.int32)
    truth = np.array([
        [24.3, 23.8, 800., math.log(2.), 0.1, -0.1],
        [32.6, 31.2, 1500., math.log(3.), -0.2, 0.05],
        [15.9, 16.4, 600., 0., 0., 0.],
        [257.5, 150.3, 5e4, math.log(8.), 0.0, 0.1],     # needs sz 1024
        [150.0, 150.0, 5e4, math.log(8.), 0.0, 0.1],
    ])
    start = truth.copy()
    start[0, :3] += (0.3, -0.2, 100.)
    start[1, :3] += (-0.4, 0.3, -200.)
    start[2, :3] += (0.2, 0.3, 50.)
    start[3, 0] = 254.5       # sz 512 now; trials beyond x = 255 need 1024
    start[4, 0] = 260.0       # its Jacobian needs 1024
    rng = np.random.default_rng(20261205)
    pix_off, total = W.patch_offsets(P)
    data = np.zeros(total, np.float32)
    for s in range(len(kind)):
        prob = O.SourceProblem(int(kind[s]), truth[s], int(P[s]), psf, cdi=B.PIXEL_CDI)
        m = O.source_render(prob, truth[s]) * np.float32(truth[s, 2])
        o = int(pix_off[s])
        data[o:o + P[s] * P[s]] = (m + rng.normal(size=m.shape)).astype(np.float32).ravel()
    return kind, P, truth, start, psf, pix_off, data


def _fit(kind, P, start, psf, pix_off, data, sel):
    offs = np.concatenate([[0], np.cumsum(P[sel].astype(np.int64) ** 2)])[:-1]
    d = np.concatenate([data[pix_off[s]:pix_off[s] + P[s] * P[s]] for s in sel])
    dd = torch.from_numpy(d).cuda()
    ie = torch.ones_like(dd)
    f = B.BatchFitter(kind[sel], P[sel], offs, dd, ie, psf)
    f.set_params(start[sel])
    f.fit(max_iter=100)
    st = f.status()
    return f.get_params(), st, f.flags()


def test_unsupported_fft_sizes_flag_sources_not_the_batch():
    assert torch.cuda.is_available()
    kind, P, truth, start, psf, pix_off, data = _scene()
    say('fit with the two edge sources')
    p_all, st, fl = _fit(kind, P, start, psf, pix_off, data, np.arange(5))
    say('flags', fl, 'iterations', st['iterations'])
    # the Jacobian source never moved and took no iteration
    assert fl[4] & TRH_FIT_F_JAC_SIZE
    assert st['iterations'][4] == 0 and st['active'][4] == 0
    assert np.array_equal(p_all[4], start[4])
    # the line-search source was flagged, moved at most to x = 255 (never to
    # a position that needs sz 1024) and ended finite
    assert fl[3] & TRH_FIT_F_LS_SIZE and not fl[3] & TRH_FIT_F_JAC_SIZE
    assert 254.5 < p_all[3, 0] <= 255.0
    assert np.isfinite(st['chisq'][3])
    # the other sources are unaffected: the same fit without the edge sources
    say('fit without them')
    p_ref, st_ref, fl_ref = _fit(kind, P, start, psf, pix_off, data, np.arange(3))
    assert np.array_equal(p_all[:3], p_ref)
    assert np.array_equal(st['iterations'][:3], st_ref['iterations'])
    assert not np.any(fl[:3] & (TRH_FIT_F_JAC_SIZE | TRH_FIT_F_LS_SIZE))


def test_max_iter_is_passed_through():
    """BatchFitter.fit(max_iter=N) stops every source after N iterations and
    flags the ones that had not met the stopping rule (optimize_loop's
    `steps`); N above 100 is honoured too."""
    assert torch.cuda.is_available()
    kind, P, truth, start, psf, pix_off, data = _scene()
    sel = np.arange(3)
    offs = np.concatenate([[0], np.cumsum(P[sel].astype(np.int64) ** 2)])[:-1]
    d = np.concatenate([data[pix_off[s]:pix_off[s] + P[s] * P[s]] for s in sel])
    dd = torch.from_numpy(d).cuda()
    f = B.BatchFitter(kind[sel], P[sel], offs, dd, torch.ones_like(dd), psf)
    for graph in (False, True):
        f.set_params(start[sel])
        f.fit(max_iter=2, graph=graph)
        st = f.status()
        fl = f.flags()
        say('max_iter 2, graph', graph, st['iterations'], fl)
        assert np.all(st['iterations'] <= 2) and st['n_active'] == 0
        capped = (fl & TRH_FIT_F_MAXITER) != 0
        assert np.any(capped) and np.all(st['iterations'][capped] == 2)
    f.set_params(start[sel])
    f.fit(max_iter=150)
    st = f.status()
    assert st['n_active'] == 0 and np.all(st['iterations'] <= 150)
