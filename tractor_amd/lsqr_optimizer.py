"""LsqrOptimizer: linearized least-squares update + line search
(tractor/lsqr_optimizer.py:253-644).

The update direction is the (column-scaled, damped) least-squares solution the
reference obtains with scipy's LSQR; here the normal equations are formed on
the device (FP32 products, FP64 sums) and solved by FP64 Cholesky
(trh_normal_eq_dense / trh_solve_normal_f64).  Column scaling, the 1e-8 drop of
tiny scaled values, prior rows and the shared-parameter map follow the
reference.
"""
import numpy as np
import torch

from .optimize import Optimizer
from .engine import logverb
from .utils import listmax
from . import _lib
from ._lib import check, ptr


def solve_least_squares(cols, b, nrows, ncols, damp=0.):
    """Least squares over sparse columns: cols[c] = (row indices, values);
    returns x minimising ||A x - b||^2 + damp^2 ||x||^2 (zero for empty
    columns), computed on the device."""
    touched = [r for (r, v) in cols if len(r)]
    if len(touched) == 0:
        return np.zeros(ncols)
    rows = np.unique(np.concatenate(touched))
    nr = len(rows)
    A = np.zeros((ncols + 1, nr), np.float32)
    for c, (r, v) in enumerate(cols):
        if len(r):
            np.add.at(A[c], np.searchsorted(rows, r), v)
    A[ncols] = b[rows]
    dev = 'cuda'
    dA = torch.from_numpy(A.ravel()).to(dev)
    nb = torch.zeros((ncols + 1) * (ncols + 1), dtype=torch.float64, device=dev)
    x = torch.zeros(ncols, dtype=torch.float64, device=dev)
    lib = _lib.load()
    h = _lib.handle(0)
    check(lib.trh_normal_eq_dense(h.value, ptr(dA), nr, ncols, ptr(nb), None), 'normal_eq_dense')
    info = np.zeros(1, np.int32)
    # solve in the (already scaled) variables: no further column scaling
    check(lib.trh_solve_normal_f64(h.value, ptr(nb), ncols, float(damp), 0, ptr(x), ptr(info), None),
          'solve_normal_f64')
    if info[0] != 0:
        # rank deficient: a tiny ridge gives the minimum-norm-like solution
        check(lib.trh_solve_normal_f64(h.value, ptr(nb), ncols, 1e-6, 0, ptr(x), ptr(info), None),
              'solve_normal_f64 (ridge)')
    return x.cpu().numpy()


class LsqrOptimizer(Optimizer):

    def getLinearUpdateDirection(self, tractor, damp=0, priors=True, scale_columns=True,
                                 shared_params=True, variance=False, **kwargs):
        allderivs = tractor.getDerivs()
        return self.getUpdateDirection(tractor, allderivs, damp=damp, priors=priors,
                                       scale_columns=scale_columns, shared_params=shared_params,
                                       variance=variance)

    def optimize(self, tractor, alphas=None, damp=0, priors=True, scale_columns=True,
                 shared_params=True, variance=False, just_variance=False, **nil):
        X = self.getLinearUpdateDirection(tractor, damp=damp, priors=priors,
                                          scale_columns=scale_columns, shared_params=shared_params,
                                          variance=variance)
        if X is None:
            return (0., None, 0.)
        if variance:
            if len(X) == 0:
                return 0, X, 0, None
            X, var = X
            if just_variance:
                return var
        if len(X) == 0:
            return 0, X, 0.
        dlogprob, alpha = self.tryUpdates(tractor, X, alphas=alphas)
        if variance:
            return dlogprob, X, alpha, var
        return dlogprob, X, alpha

    def optimize_loop(self, tractor, dchisq=0., steps=50, **kwargs):
        R = {}
        step = 0
        for step in range(steps):
            dlnp, X, alpha = self.optimize(tractor, **kwargs)
            if dlnp <= dchisq:
                break
        R.update(steps=step)
        return R

    def getUpdateDirection(self, tractor, allderivs, damp=0., priors=True, scale_columns=True,
                           scales_only=False, chiImages=None, variance=False,
                           shared_params=True, get_A_matrix=False):
        """tractor/lsqr_optimizer.py:328-644."""
        if shared_params:
            p0 = tractor.getParams()
            tractor.setParams(np.arange(len(p0)))
            p1 = tractor.getParams()
            tractor.setParams(p0)
            U, I = np.unique(p1, return_inverse=True)
            paramindexmap = I
        sprows, spcols, spvals = [], [], []
        imgoffs = {}
        nextrow = 0
        for param in allderivs:
            for deriv, img in param:
                if img in imgoffs:
                    continue
                imgoffs[img] = nextrow
                nextrow += img.numberOfPixels()
        Nrows = nextrow
        Ncols = len(allderivs)
        colscales = np.zeros(Ncols)
        for col, param in enumerate(allderivs):
            for (deriv, img) in param:
                inverrs = img.getInvError()
                H, W = img.shape
                row0 = imgoffs[img]
                deriv = deriv.copy()
                deriv.clipTo(W, H)
                pix = deriv.getPixelIndices(img)
                if len(pix) == 0:
                    continue
                dimg = deriv.getImage()
                w = inverrs[deriv.getSlice(img)]
                vals = dimg * w
                nz = np.flatnonzero(vals)
                if len(nz) == 0:
                    continue
                vals = vals.flat[nz]
                if scale_columns:
                    colscales[col] += np.dot(vals, vals)
                if scales_only:
                    continue
                sprows.append(row0 + pix[nz])
                spvals.append(vals)
                spcols.append(col)
        b = None
        if priors:
            X = tractor.getLogPriorDerivatives()
            if X is not None:
                rA, cA, vA, pb, _ = X
                for ri, ci, vi in zip(rA, cA, vA):
                    if scale_columns:
                        colscales[ci] += np.dot(vi, vi)
                    if scales_only:
                        continue
                    sprows.append(np.array(ri) + Nrows)
                    spcols.append(ci)
                    spvals.append(np.array(vi, np.float64))
                oldnrows = Nrows
                Nrows += listmax(rA, -1) + 1
                b = np.zeros(Nrows, np.float32)
                b[oldnrows:] = np.hstack(pb)
        if len(spcols) == 0:
            return []
        if scales_only:
            return np.sqrt(colscales)
        if variance:
            var = colscales.copy()
        if scale_columns:
            colscales = np.sqrt(colscales)
            rr, cc, vv = [], [], []
            for r, c, v in zip(sprows, spcols, spvals):
                scale = colscales[c]
                if scale == 0:
                    continue
                v = v / scale
                I = np.flatnonzero(np.abs(v) > 1e-8)   # drop tiny values (lsqr_optimizer.py:479-481)
                if len(I) == 0:
                    continue
                rr.append(r[I])
                vv.append(v[I])
                cc.append(c)
            sprows, spcols, spvals = rr, cc, vv
        if shared_params:
            spcols = [paramindexmap[c] for c in spcols]
            Ncols = int(np.max(spcols)) + 1 if len(spcols) else 0
        if b is None:
            b = np.zeros(Nrows, np.float32)
        chimap = {}
        if chiImages is not None:
            for img, chi in zip(tractor.getImages(), chiImages):
                chimap[img] = chi
        for img, row0 in imgoffs.items():
            chi = chimap.get(img, None)
            if chi is None:
                chi = tractor.getChiImage(img=img)
            chi = chi.ravel()
            b[row0:row0 + len(chi)] = chi
        if not np.all(np.isfinite(np.hstack(spvals))):
            print('Warning: infinite derivatives; bailing out')
            return None
        cols = [(np.zeros(0, np.int64), np.zeros(0))] * Ncols
        cols = [list(c) for c in cols]
        for r, c, v in zip(sprows, spcols, spvals):
            cr, cv = cols[c]
            cols[c] = [np.concatenate([cr, r]), np.concatenate([cv, v])]
        Xs = solve_least_squares([tuple(c) for c in cols], b.astype(np.float64), Nrows, Ncols,
                                 damp=damp)
        if shared_params:
            Xs = Xs[paramindexmap]
        X = np.array(Xs, np.float64)
        if scale_columns:
            X[colscales > 0] /= colscales[colscales > 0]
        if variance:
            if shared_params:
                var = var[paramindexmap]
            return X, 1. / np.array(var)
        return X
