"""Pins the oracle's line search and stopping rule with hand-computed cases
(ConstrainedOptimizer.tryUpdates / getParameterSteps / optimize_loop,
tractor/constrained_optimizer.py:12-159).

Each case is built so that a plausible misreading of the rule gives a
different answer; the mutation checks at the end run those misreadings on
the same cases and require at least one case to catch each of them."""
import math

import numpy as np
import pytest

import oracle as O

A = O.ALPHAS


def _steps(lnps):
    # (alpha, p, step_limit, hit_limit) records, one per trial
    return [(float(A[i]), [float(A[i])], False, i == 1) for i in range(len(lnps))]


# (lnp0, trial lnps, expected index of the accepted trial or None, expected dlnp)
CASES = [
    # best of several: the first improving trial (0) is not the best; 2.5 at
    # index 3 is below the best 3 but within 1 of it, so the walk goes on and
    # index 4 (3.5) wins
    (0., [1., 3., 2., 2.5, 3.5], 4, 3.5),
    # a drop of more than 1 below the best ends the walk: the 10 after it is
    # never reached
    (0., [1., 2., 0.9, 10.], 1, 2.),
    # exactly 1 below the best does not end it (the test is lnp < best - 1)
    (0., [1., 2., 1., 10.], 3, 10.),
    # a non-finite lnp ends the walk
    (0., [1., float('nan'), 10.], 0, 1.),
    (0., [1., -math.inf, 10.], 0, 1.),
    (0., [1., math.inf, 10.], 0, 1.),
    # nothing improves: no step, dlnp 0
    (5., [4.5, 4.9, 4.2], None, 0.),
    # equal to the start is not an improvement (strict >)
    (5., [5., 5.], None, 0.),
    # the break compares with the best so far, not the start: 3 is 2 below
    # the best 5 (but above the start)
    (0., [5., 3., 9.], 0, 5.),
]


@pytest.mark.parametrize('lnp0,lnps,ib,dlnp', CASES)
def test_accept_line_search_hand_cases(lnp0, lnps, ib, dlnp):
    steps = _steps(lnps)
    got_ib, got_dlnp, alpha, hl = O.accept_line_search(lnp0, steps, lnps)
    assert got_ib == ib
    assert got_dlnp == pytest.approx(dlnp, abs=0)
    if ib is None:
        assert alpha == 0. and hl is False
    else:
        assert alpha == A[ib]
        # last_step_hit_limit is the accepted trial's hit_limit
        assert hl == (ib == 1)


def test_parameter_steps_flags():
    """stepLimited is set by any max-step clamp (constrained_optimizer.py:
    113-121), kept even when the clamped alpha then falls below tiny_alpha
    and the trial is dropped."""
    # no limit reached: 13 trials, not step-limited
    fl = {'step_limited': False}
    st = O.parameter_steps([0.], [1e-3], [1.], [None], [None], flags=fl)
    assert len(st) == 13 and not fl['step_limited']
    # clamp at alpha = 1 / 3 (|X| = 3, maxstep 1): 2^-10 .. 2^-2 pass,
    # 2^-1 is clamped to 1/3 and is the last trial
    fl = {'step_limited': False}
    st = O.parameter_steps([0.], [3.], [1.], [None], [None], flags=fl)
    assert fl['step_limited']
    assert [s[0] for s in st] == list(A[:9]) + [1. / 3.]
    assert [s[2] for s in st] == [False] * 9 + [True]
    assert st[-1][1] == [1.]
    # clamped alpha below tiny_alpha: nothing to try, still step-limited
    fl = {'step_limited': False}
    st = O.parameter_steps([0.], [1e9], [1.], [None], [None], flags=fl)
    assert st == [] and fl['step_limited']
    # a bound clamps alpha (hit_limit) but is not a step limit
    fl = {'step_limited': False}
    st = O.parameter_steps([0.5], [1.], [None], [0.], [0.6], flags=fl)
    assert not fl['step_limited']
    assert st[-1][3] is True and st[-1][1] == [0.6]
    assert st[-1][0] == pytest.approx(0.1)
    # a bound right at the start: alpha 0 < tiny_alpha, no trial at all
    fl = {'step_limited': False}
    st = O.parameter_steps([0.6], [1.], [None], [0.], [0.6], flags=fl)
    assert st == [] and not fl['step_limited']


STOP_CASES = [
    # (step_limited, dlnp, stop?) at dchisq = 0, dchisq_limited = 1e-6
    (False, 0., True),
    (False, -1., True),
    (False, 1e-9, False),
    (True, 1e-9, True),      # a limited step must gain more than 1e-6
    (True, 1e-6, True),
    (True, 2e-6, False),
    (True, 0., True),
]


@pytest.mark.parametrize('limited,dlnp,stop', STOP_CASES)
def test_loop_should_stop(limited, dlnp, stop):
    assert O.loop_should_stop(limited, dlnp, 0., 1e-6) is stop


class _Quadratic(object):
    """A 1-parameter stand-in for SourceProblem whose chi^2 is a parabola, so
    the optimizer's path can be followed by hand: chi^2(p) = 4 (p - 3)^2 and
    the update direction is the exact Newton step 3 - p."""
    def __init__(self, maxstep):
        self.ms = maxstep


def _quad_fit(p0, maxstep, steps=100):
    """fit_source's loop on _Quadratic (same helpers, same order)."""
    chi = lambda p: 4. * (p - 3.) ** 2
    params = p0
    it = 0
    trace = []
    for it in range(steps):
        X = [3. - params]
        lnp0 = -0.5 * chi(params)
        fl = {'step_limited': False}
        stps = O.parameter_steps([params], X, [maxstep], [None], [None], flags=fl)
        lim = fl['step_limited'] or len(stps) == 0
        lnps = [-0.5 * chi(s[1][0]) for s in stps]
        ib, dlnp, alpha, hl = O.accept_line_search(lnp0, stps, lnps)
        if ib is not None:
            params = stps[ib][1][0]
        trace.append((alpha, lim))
        if O.loop_should_stop(lim, dlnp, 0., 1e-6):
            break
    return params, it + 1, trace


def test_loop_on_a_parabola():
    # from p = 0 with maxstep 1:
    #  1. X = 3: alphas up to 1/4 pass (step 0.75), 1/2 is clamped to 1/3
    #     (p = 1, the best trial); step-limited.
    #  2. X = 2: alpha 1/2 reaches p = 2 (step exactly 1, not clamped), alpha 1
    #     is clamped back to 1/2 (p = 2 again, not better): alpha 1/2, limited.
    #  3. X = 1: alpha 1 lands on 3; sqrt 2 is clamped to 1 (same p, not
    #     better): alpha 1, still step-limited (the clamp sets the flag even
    #     though the best trial was not clamped), dlnp = 2 > 1e-6: go on.
    #  4. X = 0: no trial improves, dlnp = 0, not limited: stop.
    p, nit, trace = _quad_fit(0., 1.)
    assert p == 3.
    assert nit == 4
    assert [t[1] for t in trace] == [True, True, True, False]
    assert trace[0][0] == pytest.approx(1. / 3.)
    assert trace[1][0] == 0.5
    assert trace[2][0] == 1.
    assert trace[3][0] == 0.


# ---------------------------------------------------------------------------
# Mutation checks: each plausible misreading, run on the cases above, must
# fail at least one of them.
# ---------------------------------------------------------------------------
def _no_drop_break(lnp0, steps, lnps):
    best, ib = lnp0, None
    for i, lnp in enumerate(lnps):
        if not np.isfinite(lnp):
            break
        if lnp > best:
            best, ib = lnp, i
    return ib, (best - lnp0) if ib is not None else 0.


def _first_improving(lnp0, steps, lnps):
    for i, lnp in enumerate(lnps):
        if not np.isfinite(lnp):
            break
        if lnp > lnp0:
            return i, lnp - lnp0
    return None, 0.


def _break_vs_start(lnp0, steps, lnps):
    best, ib = lnp0, None
    for i, lnp in enumerate(lnps):
        if not np.isfinite(lnp) or lnp < lnp0 - 1.:
            break
        if lnp > best:
            best, ib = lnp, i
    return ib, (best - lnp0) if ib is not None else 0.


def _nonstrict(lnp0, steps, lnps):
    best, ib = lnp0, None
    for i, lnp in enumerate(lnps):
        if not np.isfinite(lnp) or lnp < best - 1.:
            break
        if lnp >= best:
            best, ib = lnp, i
    return ib, (best - lnp0) if ib is not None else 0.


def _skip_nonfinite(lnp0, steps, lnps):
    best, ib = lnp0, None
    for i, lnp in enumerate(lnps):
        if not np.isfinite(lnp):
            continue
        if lnp < best - 1.:
            break
        if lnp > best:
            best, ib = lnp, i
    return ib, (best - lnp0) if ib is not None else 0.


@pytest.mark.parametrize('mutant', [_no_drop_break, _first_improving, _break_vs_start,
                                    _nonstrict, _skip_nonfinite])
def test_cases_catch_misreadings(mutant):
    caught = False
    for lnp0, lnps, ib, dlnp in CASES:
        got_ib, got_dlnp = mutant(lnp0, _steps(lnps), lnps)
        if got_ib != ib or got_dlnp != dlnp:
            caught = True
    assert caught


@pytest.mark.parametrize('mutant', [
    lambda lim, d: d <= 0.,                          # dchisq_limited ignored
    lambda lim, d: d <= (0. if lim else 1e-6),       # thresholds swapped
    lambda lim, d: d <= 1e-6,                        # one threshold for both
])
def test_stop_cases_catch_misreadings(mutant):
    assert any(mutant(lim, d) is not stop for lim, d, stop in STOP_CASES)
