"""The converged-fit parity rule of the batched-fit GPU tests.

A source passes `tight` when its chi^2 is within 1e-4 relative of the FP64
oracle's and every parameter within 1e-4 relative or 0.01 sigma (sigma from
the Fisher matrix at the oracle's solution).  The one named exception,
`flat`: chi^2 within 2e-6 relative and every parameter within 0.05 sigma.
There both fits stopped on the flat bottom of the same minimum: a parameter
k sigma off the minimum raises chi^2 by about k^2, and the FP32 renders
resolve chi^2 to about 1e-7 relative (a few 1e-4 absolute at chi^2 ~ 4000),
so the line search's dlnp <= 0 stop cannot place a parameter closer than a
few hundredths of sigma.  Anything else is a miss."""
import numpy as np


def classify(got, ref, sig, chi_gpu, chi_ref):
    err = np.abs(np.asarray(got) - np.asarray(ref))
    ref = np.abs(np.asarray(ref))
    rel = abs(chi_gpu - chi_ref) / chi_ref
    if rel <= 1e-4 and np.all((err <= 1e-4 * ref) | (err <= 0.01 * sig)):
        return 'tight'
    if rel <= 2e-6 and np.all((err <= 1e-4 * ref) | (err <= 0.05 * sig)):
        return 'flat'
    return 'miss'


def check(results, say, max_flat_frac=0.1):
    """results: [(source, classify(...), detail string)].  No misses; at
    most max_flat_frac of the sources are flat-bottom exceptions."""
    n = len(results)
    tight = sum(r[1] == 'tight' for r in results)
    flat = [r for r in results if r[1] == 'flat']
    miss = [r for r in results if r[1] == 'miss']
    for r in flat:
        say('  flat-bottom exception: source %d %s' % (r[0], r[2]))
    for r in miss:
        say('  MISS: source %d %s' % (r[0], r[2]))
    say('sources within 1e-4 rel or 0.01 sigma: %d / %d; flat-bottom exceptions %d' % (
        tight, n, len(flat)))
    assert not miss
    assert len(flat) <= max_flat_frac * n
