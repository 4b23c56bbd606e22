"""ConstrainedOptimizer (bounds + max steps in the line search) and the
GPU-plugin counterpart (tractor/constrained_optimizer.py:1-159,
tractor/gpu_optimizer.py:24-92)."""
import numpy as np

from .lsqr_optimizer import LsqrOptimizer
from .optimize import ALPHAS


class ConstrainedOptimizer(LsqrOptimizer):

    def __init__(self, *args, **kwargs):
        super(ConstrainedOptimizer, self).__init__(*args, **kwargs)
        self.stepLimited = False
        self.tiny_alpha = 1e-8

    def optimize_loop(self, tractor, dchisq=0., steps=100, dchisq_limited=1e-6, check_step=None,
                      **kwargs):
        R = {}
        self.hit_limit = False
        self.last_step_hit_limit = False
        step = 0
        for step in range(steps):
            self.stepLimited = False
            dlnp, X, alpha = self.optimize(tractor, **kwargs)
            if check_step is not None:
                if not check_step(optimizer=self, tractor=tractor, dlnp=dlnp, X=X, alpha=alpha):
                    break
            if not self.stepLimited and dlnp <= dchisq:
                break
            if self.stepLimited and dlnp <= dchisq_limited:
                break
        R.update(steps=step, hit_limit=self.last_step_hit_limit, ever_hit_limit=self.hit_limit)
        return R

    def tryUpdates(self, tractor, X, alphas=None, check_step=None):
        p_best = tractor.getParams()
        logprob_before = tractor.getLogProb()
        logprob_best = logprob_before
        alpha_best = None
        steps = self.getParameterSteps(tractor, X, alphas)
        if len(steps) == 0:
            self.stepLimited = True
            self.last_step_hit_limit = True
            self.hit_limit = True
        for alpha, p, step_limit, hit_limit in steps:
            if hit_limit:
                self.hit_limit = True
            tractor.setParams(p)
            if check_step is not None:
                if not check_step(optimizer=self, tractor=tractor, X=X, alpha=alpha):
                    break
            logprob = tractor.getLogProb()
            if not np.isfinite(logprob):
                break
            if logprob < (logprob_best - 1.):
                break
            if logprob > logprob_best:
                self.last_step_hit_limit = hit_limit
                alpha_best = alpha
                logprob_best = logprob
                p_best = p
        tractor.setParams(p_best)
        if alpha_best is None:
            return 0., 0.
        return logprob_best - logprob_before, alpha_best

    def getParameterSteps(self, tractor, step_direction, alphas):
        """constrained_optimizer.py:84-159."""
        if alphas is None:
            alphas = ALPHAS
        p0 = tractor.getParams()
        max_step_size = tractor.getMaxStep()
        lowers = tractor.getLowerBounds()
        uppers = tractor.getUpperBounds()
        assert np.all(np.diff(alphas) > 0)
        step_direction = np.array(step_direction)
        results = []
        for alpha in alphas:
            do_break = False
            step_limit = False
            for i, mx in enumerate(max_step_size):
                if mx is None:
                    continue
                step = alpha * abs(step_direction[i])
                if step > mx:
                    step_limit = True
                    do_break = True
                    alpha = mx / abs(step_direction[i])
                    self.stepLimited = True
            hit_limit = False
            for i, (p_start, l, u) in enumerate(zip(p0, lowers, uppers)):
                px = p_start + alpha * step_direction[i]
                if l is not None and px < l:
                    alpha = (l - p_start) / step_direction[i]
                    hit_limit = True
                if u is not None and px > u:
                    alpha = (u - p_start) / step_direction[i]
                    hit_limit = True
            if alpha < self.tiny_alpha:
                break
            p = []
            for p_start, s, l, u in zip(p0, step_direction, lowers, uppers):
                px = p_start + alpha * s
                if l is not None and px < l:
                    px = l
                if u is not None and px > u:
                    px = u
                p.append(px)
            results.append((alpha, p, step_limit, hit_limit))
            if do_break:
                break
        return results


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
