"""The oracle's C++ FP64 fit (oracle/fit_fp64.cpp) against the NumPy oracle's
fit_source on the same sources, data and start: both restate the same
reference loop with the same FP32 / FP64 split, so they differ by rounding
only (their FFTs, sums and least-squares solvers differ).  A converged source
ends at the same chi^2 within 1e-6 relative (typically 1e-10 or better; the
dchisq = 0 rule can stop one iteration apart); the C++ build is made here if
it is missing (g++, no GPU code).  Also: thread count and time budget do not
change results, and unsupported kinds are rejected."""
import os

import numpy as np
import pytest

import oracle as O
from oracle import fit_fp64 as F
from tractor_amd import workloads as W


@pytest.fixture(scope='module')
def scene():
    if not os.path.exists(F.LIB_PATH):
        F.build()
    rng = np.random.default_rng(5)
    kind, P, truth = W.draw_sources(12, rng=rng, P_choices=np.array([24, 40, 64]))
    psf = W.moffat_stamp(63, 4.0, 3.5)
    off = np.zeros(len(P), np.int64)
    off[1:] = np.cumsum(P.astype(np.int64) ** 2)[:-1]
    npx = int((P.astype(np.int64) ** 2).sum())
    data = np.zeros(npx, np.float32)
    inv = np.ones(npx, np.float32)
    nr = np.random.default_rng(9)
    for s in range(len(P)):
        Pz = int(P[s])
        prob = O.SourceProblem(int(kind[s]), truth[s], Pz, psf)
        m = (O.source_render(prob, truth[s]) * np.float32(truth[s, 2])).astype(np.float32)
        data[off[s]:off[s] + Pz * Pz] = (m + nr.normal(0, 1, (Pz, Pz))).ravel()
    start = W.perturb(kind, truth, np.random.default_rng(3))
    return dict(kind=kind, P=P, off=off, data=data, inv=inv, start=start, psf=psf)


def _numpy_fit(sc, s, **kw):
    Pz, o = int(sc['P'][s]), int(sc['off'][s])
    prob = O.SourceProblem(int(sc['kind'][s]), sc['start'][s], Pz, sc['psf'],
                           data=sc['data'][o:o + Pz * Pz].reshape(Pz, Pz),
                           inverr=sc['inv'][o:o + Pz * Pz].reshape(Pz, Pz))
    return O.fit_source(prob, sc['start'][s], **kw)


def test_cpp_fit_matches_numpy_oracle(scene):
    sc = scene
    done, p, it, chi = F.fit_sources(sc['kind'], sc['P'], sc['off'], sc['data'], sc['inv'],
                                     sc['start'], sc['psf'], nthreads=4)
    assert done == len(sc['kind'])
    for s in range(done):
        pn, nit, chin = _numpy_fit(sc, s)
        assert it[s] > 0
        if nit == 100:
            continue   # not converged: the trajectories may part
        assert abs(chi[s] - chin) <= 1e-6 * chin, (s, chi[s], chin)
        assert abs(it[s] - nit) <= 2, (s, it[s], nit)


def test_one_iteration_is_the_numpy_step(scene):
    """max_iter 1: one update and line search, the same step as the NumPy
    oracle's single iteration (no stopping rule involved)."""
    sc = scene
    done, p, it, chi = F.fit_sources(sc['kind'], sc['P'], sc['off'], sc['data'], sc['inv'],
                                     sc['start'], sc['psf'], max_iter=1, nthreads=2)
    for s in range(done):
        pn, nit, chin = _numpy_fit(sc, s, steps=1)
        k = 3 if sc['kind'][s] == 0 else 6
        assert it[s] == 1
        np.testing.assert_allclose(p[s, :k], pn[:k], rtol=1e-7, atol=1e-9)
        assert abs(chi[s] - chin) <= 1e-9 * chin


def test_threads_and_budget_do_not_change_results(scene):
    sc = scene
    a = F.fit_sources(sc['kind'], sc['P'], sc['off'], sc['data'], sc['inv'], sc['start'],
                      sc['psf'], nthreads=1)
    b = F.fit_sources(sc['kind'], sc['P'], sc['off'], sc['data'], sc['inv'], sc['start'],
                      sc['psf'], nthreads=8, budget_s=1e-9)
    # a tiny budget still finishes the sources its threads had started
    assert 1 <= b[0] <= 8
    np.testing.assert_array_equal(a[1][:b[0]], b[1][:b[0]])
    np.testing.assert_array_equal(a[3][:b[0]], b[3][:b[0]])


def test_rejects_unsupported_kind(scene):
    sc = scene
    k = sc['kind'].copy()
    k[0] = 3
    with pytest.raises(ValueError):
        F.fit_sources(k, sc['P'], sc['off'], sc['data'], sc['inv'], sc['start'], sc['psf'])
