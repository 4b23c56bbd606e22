"""GPU parity of the fitter's stages as standalone batched exports
(trh_jacobian_normal_eq_batch, trh_solve_damped_cholesky_f64_batch,
trh_line_search_chi2_batch) against the FP64 oracle, on the fitter's own
renders and data.

Tolerances: the normal equations within 3e-3 relative of their largest
entry (the FP32 forward differences of the shape columns amplify render
rounding ~100x, as in the fitter's own test); the step within 3e-3 of the
oracle's column-scaled least-squares step; chi^2 of trial renders within
1e-5 relative (FP32 render, FP64 sums)."""
import ctypes
import sys

import numpy as np
import pytest
import torch

import oracle as O
from tractor_amd import _lib, batch as B, workloads as W
from tractor_amd._lib import check, ptr

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope='module')
def pb():
    assert torch.cuda.is_available()
    rng = np.random.default_rng(W.seed(2, 20))
    kind, P, truth = W.draw_sources(24, rng=rng, P_choices=np.array([24, 40, 64, 88, 120]))
    psf = W.moffat_stamp(63, 4.0, 3.5)
    f, data, inverr, pix_off = B.make_problem(kind, P, truth, psf, W.seed(2, 21))
    start = W.perturb(kind, truth, np.random.default_rng(W.seed(2, 22)))
    return dict(kind=kind, P=P, psf=psf, f=f, data=data, inverr=inverr, pix_off=pix_off,
                start=start, dh=data.cpu().numpy(), ih=inverr.cpu().numpy())


def _prob(pb, s, p):
    Pz, o = int(pb['P'][s]), int(pb['pix_off'][s])
    return O.SourceProblem(int(pb['kind'][s]), p, Pz, pb['psf'],
                           data=pb['dh'][o:o + Pz * Pz].reshape(Pz, Pz),
                           inverr=pb['ih'][o:o + Pz * Pz].reshape(Pz, Pz), cdi=B.PIXEL_CDI)


def test_jacobian_normal_eq_and_solve_exports(pb):
    say('normal equations + solve exports')
    f = pb['f']
    S = len(pb['kind'])
    f.set_params(pb['start'])
    f.iterate(0., 0.)                       # fills the fitter's Jacobian renders
    lib = _lib.load()
    h = _lib.handle(0).value
    r, ro = ctypes.c_void_p(), ctypes.c_void_p()
    check(lib.trh_fitter_debug(f.f, ctypes.byref(r), ctypes.byref(ro), None, None), 'debug')
    dev = torch.device('cuda')
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    d_kind, d_P = t(pb['kind'].astype(np.int32)), t(pb['P'].astype(np.int32))
    d_stride = t(pb['P'].astype(np.int32))
    d_off = t(pb['pix_off'].astype(np.int64))
    d_counts = t(pb['start'][:, 2].astype(np.float64))
    neq = torch.full((S, 40), float('nan'), dtype=torch.float64, device=dev)
    cdi = np.eye(2).ravel()
    check(lib.trh_jacobian_normal_eq_batch(h, S, ptr(d_kind), ptr(d_P), ptr(d_stride), ptr(d_off),
                                           ptr(pb['data']), ptr(pb['inverr']), ro.value, r.value,
                                           ptr(d_counts), None, ptr(cdi), 0.01, ptr(neq), None),
          'trh_jacobian_normal_eq_batch')
    X = torch.zeros((S, 8), dtype=torch.float64, device=dev)
    check(lib.trh_solve_damped_cholesky_f64_batch(h, S, ptr(d_kind), ptr(neq), 0., None, ptr(X),
                                                  None, None, None, None),
          'trh_solve_damped_cholesky_f64_batch')
    torch.cuda.synchronize()
    neq, X = neq.cpu().numpy(), X.cpu().numpy()
    # the fitter's own normal equations of the same iteration: same kernel
    assert np.array_equal(neq[:, :38], f.debug_arrays()['neq'][:, :38])
    for s in range(S):
        prob = _prob(pb, s, pb['start'][s])
        p0, derivs = O.source_jacobian(prob, pb['start'][s])
        A = np.stack([(d * prob.inverr).astype(np.float32).ravel() for d in derivs],
                     1).astype(np.float64)
        mod = (p0 * np.float32(pb['start'][s, 2])).astype(np.float32)
        b = ((prob.data - mod) * prob.inverr).astype(np.float32).ravel().astype(np.float64)
        nc = A.shape[1]
        N = A.T @ A
        q = neq[s]
        got = np.zeros((nc, nc))
        k = 0
        for i in range(6):
            for j in range(i, 6):
                if i < nc and j < nc:
                    got[i, j] = got[j, i] = q[k]
                k += 1
        assert np.max(np.abs(got - N)) <= 3e-3 * np.max(np.abs(N)), s
        assert np.max(np.abs(q[21:21 + nc] - A.T @ b)) <= 3e-3 * np.max(np.abs(A.T @ b)) + 1e-9
        assert abs(q[36] - b @ b) <= 1e-5 * (b @ b)
        Xo = O.column_scaled_lstsq(A, b)
        sc = np.sqrt(np.diag(N))
        assert np.max(np.abs((X[s, :nc] - Xo) * sc)) <= 3e-3 * np.max(np.abs(Xo * sc)), s


def test_line_search_chi2_export(pb):
    say('line-search chi2 export')
    f = pb['f']
    S = len(pb['kind'])
    rng = np.random.default_rng(7)
    trial = np.zeros((S, 13, 8))
    nsteps = rng.integers(1, 14, S).astype(np.int32)
    for s in range(S):
        for a in range(13):
            p = pb['start'][s].copy()
            p[:2] += rng.uniform(-0.5, 0.5, 2)
            p[2] *= rng.uniform(0.8, 1.2)
            if pb['kind'][s]:
                p[3:6] += rng.uniform(-0.05, 0.05, 3)
            trial[s, a, :6] = p
    dev = torch.device('cuda')
    d_trial = torch.from_numpy(trial).to(dev)
    d_n = torch.from_numpy(nsteps).to(dev)
    chi2 = torch.full((S, 13), float('nan'), dtype=torch.float64, device=dev)
    lib = _lib.load()
    check(lib.trh_line_search_chi2_batch(f.f, ptr(d_trial), ptr(d_n), ptr(chi2), None),
          'trh_line_search_chi2_batch')
    torch.cuda.synchronize()
    chi2 = chi2.cpu().numpy()
    worst = 0.
    for s in range(S):
        for a in range(13):
            if a >= nsteps[s]:
                continue
            ref = O.source_chisq(_prob(pb, s, trial[s, a, :6]), trial[s, a, :6])
            e = abs(chi2[s, a] - ref) / ref
            worst = max(worst, e)
            assert e <= 1e-5, (s, a, chi2[s, a], ref)
    say('  worst', worst)
