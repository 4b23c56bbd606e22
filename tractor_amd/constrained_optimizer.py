"""ConstrainedOptimizer (bounds and maximum steps in the line search) and the
GPU-plugin counterpart (tractor/constrained_optimizer.py:1-159,
tractor/gpu_optimizer.py:24-92).

The line search evaluates every trial step's log-probability in one
RenderPlan (Tractor._log_probs) and then walks them in order with the
reference's acceptance rule; the reference renders them one at a time and
stops early, which picks the same step (the walk stops where it would).
"""
import numpy as np

from .lsqr_optimizer import LsqrOptimizer
from .optimize import ALPHAS, walk_line_search


def _within(v, lo, hi):
    if lo is not None and v < lo:
        v = lo
    if hi is not None and v > hi:
        v = hi
    return v


class ConstrainedOptimizer(LsqrOptimizer):

    def __init__(self, *args, **kwargs):
        super(ConstrainedOptimizer, self).__init__(*args, **kwargs)
        self.stepLimited = False
        self.tiny_alpha = 1e-8

    def optimize_loop(self, tractor, dchisq=0., steps=100, dchisq_limited=1e-6, check_step=None,
                      **kwargs):
        """Steps until one gains at most dchisq, or at most dchisq_limited
        when the step was limited by a maximum step size
        (constrained_optimizer.py:12-34).  Returns the last step's index and
        whether the last accepted / any step reached a bound."""
        self.hit_limit = False
        self.last_step_hit_limit = False
        k = 0
        for k in range(steps):
            self.stepLimited = False
            dlnp, X, alpha = self.optimize(tractor, **kwargs)
            if check_step is not None and not check_step(optimizer=self, tractor=tractor,
                                                         dlnp=dlnp, X=X, alpha=alpha):
                break
            if dlnp <= (dchisq_limited if self.stepLimited else dchisq):
                break
        return dict(steps=k, hit_limit=self.last_step_hit_limit, ever_hit_limit=self.hit_limit)

    def tryUpdates(self, tractor, X, alphas=None, check_step=None):
        """The constrained line search (constrained_optimizer.py:36-82) ->
        (dlnp, alpha); leaves the tractor at the accepted step (or where it
        started)."""
        p_start = tractor.getParams()
        steps = self.getParameterSteps(tractor, X, alphas)
        if not steps:
            self.stepLimited = True
            self.last_step_hit_limit = True
            self.hit_limit = True
        def visit(i):
            # a step that reached a bound marks the search once it is tried
            if steps[i][3]:
                self.hit_limit = True
        if check_step is None:
            lnps = tractor._log_probs([None] + [p for (_, p, _, _) in steps])
            lnp0, trial = lnps[0], lnps[1:]
        else:
            # a callback may end the search before a step is evaluated:
            # evaluate them one at a time
            lnp0 = tractor.getLogProb()
            trial = []
            for i, (alpha, p, _, _) in enumerate(steps):
                visit(i)
                tractor.setParams(p)
                if not check_step(optimizer=self, tractor=tractor, X=X, alpha=alpha):
                    break
                trial.append(tractor.getLogProb())
        best, lnp = walk_line_search(lnp0, trial, visit if check_step is None else None)
        if best is None:
            tractor.setParams(p_start)
            return 0., 0.
        alpha, p, _, hit = steps[best]
        self.last_step_hit_limit = hit
        tractor.setParams(p)
        return lnp - lnp0, alpha

    def getParameterSteps(self, tractor, step_direction, alphas):
        """[(alpha, parameters, step_limit, hit_limit)] for each alpha in
        increasing order (constrained_optimizer.py:84-159): alpha is cut to
        the largest value whose step stays within every parameter's maximum
        step (step_limit; the last trial) and then to where the step meets
        a bound (hit_limit); a cut below tiny_alpha ends the list."""
        alphas = ALPHAS if alphas is None else alphas
        assert np.all(np.diff(alphas) > 0)
        p0 = [float(v) for v in tractor.getParams()]
        X = [float(v) for v in step_direction]
        maxstep = tractor.getMaxStep()
        bounds = list(zip(tractor.getLowerBounds(), tractor.getUpperBounds()))
        out = []
        for a in alphas:
            limited = False
            for mx, x in zip(maxstep, X):
                if mx is not None and a * abs(x) > mx:
                    a, limited = mx / abs(x), True
            if limited:
                self.stepLimited = True
            hit = False
            for p, x, (lo, hi) in zip(p0, X, bounds):
                v = p + a * x
                if lo is not None and v < lo:
                    a, v, hit = (lo - p) / x, lo, True
                if hi is not None and v > hi:
                    a, hit = (hi - p) / x, True
            if a < self.tiny_alpha:
                break
            out.append((a, [_within(p + a * x, lo, hi) for p, x, (lo, hi) in zip(p0, X, bounds)],
                        limited, hit))
            if limited:
                break
        return out


class GpuOptimizer(ConstrainedOptimizer):
    """The reference's GPU plugin interface, GpuOptimizer(cp)
    (tractor/gpu_optimizer.py:63-92): dense update without shared parameters
    and the constrained line search, all renders and the least-squares step on
    the MI355X.  `cp` is accepted for interface compatibility ('hip', 'cupy'
    and 'numpy' all select the HIP backend; there is no CPU backend).  For
    thousands of one-source fits use tractor_amd.batch.BatchFitter."""

    def __init__(self, cp='hip', *args, **kwargs):
        super(GpuOptimizer, self).__init__(*args, **kwargs)
        self.cp_orig = cp

    def __repr__(self):
        return 'GpuOptimizer(hip)'

    def optimize(self, tractor, alphas=None, damp=0, priors=True, scale_columns=True,
                 shared_params=False, variance=False, just_variance=False, **nil):
        return super(GpuOptimizer, self).optimize(tractor, alphas=alphas, damp=damp, priors=priors,
                                                  scale_columns=scale_columns,
                                                  shared_params=shared_params, variance=variance,
                                                  just_variance=just_variance)
