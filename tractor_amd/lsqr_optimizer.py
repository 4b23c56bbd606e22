"""LsqrOptimizer: linearized least-squares update + line search
(tractor/lsqr_optimizer.py:253-644).

The update direction is the least-squares solution the reference obtains
with scipy's LSQR on its sparse, column-scaled system; here the normal
equations of that system are formed on the device (FP32 products, FP64
sums, trh_normal_eq_dense) and solved by FP64 Cholesky
(trh_solve_normal_f64).  The system's rules are the reference's: one row
per image pixel (images stacked in the order their derivatives first
appear) and one per prior term; weights inverr; columns scaled to unit norm
and entries at most 1e-8 dropped after scaling; duplicate parameters merged.
"""
import numpy as np
import torch

from . import _lib
from ._lib import check, ptr
from .optimize import Optimizer


def solve_least_squares(cols, b, nrows, ncols, damp=0.):
    """min ||A x - b||^2 + damp^2 ||x||^2 over sparse columns cols[c] =
    (row indices, values), on the device (zero for empty columns)."""
    touched = [r for (r, v) in cols if len(r)]
    if not touched:
        return np.zeros(ncols)
    rows = np.unique(np.concatenate(touched))
    A = np.zeros((ncols + 1, len(rows)), np.float32)
    for c, (r, v) in enumerate(cols):
        if len(r):
            # entries of one column may repeat a row (several patches): sum them
            A[c] += np.bincount(np.searchsorted(rows, r), weights=v,
                                minlength=len(rows)).astype(np.float32)
    A[ncols] = b[rows]
    dev = 'cuda'
    dA = torch.from_numpy(A.ravel()).to(dev)
    nb = torch.zeros((ncols + 1) ** 2, dtype=torch.float64, device=dev)
    x = torch.zeros(ncols, dtype=torch.float64, device=dev)
    lib = _lib.load()
    h = _lib.handle(0).value
    check(lib.trh_normal_eq_dense(h, ptr(dA), len(rows), ncols, ptr(nb), None), 'normal_eq_dense')
    info = np.zeros(1, np.int32)
    # the columns are already scaled: no further scaling in the solve
    check(lib.trh_solve_normal_f64(h, ptr(nb), ncols, float(damp), 0, ptr(x), ptr(info), None),
          'solve_normal_f64')
    if info[0] != 0:
        # rank deficient: a tiny ridge picks a finite solution
        check(lib.trh_solve_normal_f64(h, ptr(nb), ncols, 1e-6, 0, ptr(x), ptr(info), None),
              'solve_normal_f64 (ridge)')
    return x.cpu().numpy()


class _SparseSystem(object):
    """The weighted design matrix of one linearized step, column by column."""

    def __init__(self, ncols):
        self.first_row = {}       # image -> its first row
        self.nrows = 0
        self.rows = [[] for _ in range(ncols)]
        self.vals = [[] for _ in range(ncols)]
        self.sumsq = np.zeros(ncols)

    def image_row(self, img):
        r = self.first_row.get(img)
        if r is None:
            r = self.first_row[img] = self.nrows
            self.nrows += img.numberOfPixels()
        return r

    def add(self, col, rows, vals, keep=True):
        self.sumsq[col] += np.dot(vals, vals)
        if keep:
            self.rows[col].append(rows)
            self.vals[col].append(vals)

    def add_patch(self, col, deriv, img, keep=True):
        """deriv x inverr on the image pixels it covers, nonzero entries only."""
        H, W = img.shape
        d = deriv.copy()
        d.clipTo(W, H)
        pix = d.getPixelIndices(img)
        if len(pix) == 0:
            return
        v = (d.getImage() * img.getInvError()[d.getSlice(img)]).ravel()
        nz = np.flatnonzero(v)
        if len(nz):
            self.add(col, self.image_row(img) + pix[nz], v[nz], keep)

    def entries(self):
        for c in range(len(self.rows)):
            for r, v in zip(self.rows[c], self.vals[c]):
                yield c, r, v


class LsqrOptimizer(Optimizer):

    def getLinearUpdateDirection(self, tractor, damp=0, priors=True, scale_columns=True,
                                 shared_params=True, variance=False, **kwargs):
        return self.getUpdateDirection(tractor, tractor.getDerivs(), damp=damp, priors=priors,
                                       scale_columns=scale_columns, shared_params=shared_params,
                                       variance=variance)

    def optimize(self, tractor, alphas=None, damp=0, priors=True, scale_columns=True,
                 shared_params=True, variance=False, just_variance=False, **nil):
        """(dlnp, X, alpha), with the variances appended when `variance`
        (only them with just_variance); (0, None, 0) when no direction can
        be found (lsqr_optimizer.py:280-315)."""
        got = self.getLinearUpdateDirection(tractor, damp=damp, priors=priors,
                                            scale_columns=scale_columns,
                                            shared_params=shared_params, variance=variance)
        if got is None:
            return (0., None, 0.)
        var = None
        if variance:
            if len(got) == 0:
                return 0, got, 0, None
            got, var = got
            if just_variance:
                return var
        if len(got) == 0:
            return (0, got, 0.) if not variance else (0, got, 0., var)
        dlnp, alpha = self.tryUpdates(tractor, got, alphas=alphas)
        return (dlnp, got, alpha) if not variance else (dlnp, got, alpha, var)

    def optimize_loop(self, tractor, dchisq=0., steps=50, **kwargs):
        """Steps until one gains at most dchisq (lsqr_optimizer.py:317-326);
        'steps' is the index of the last one."""
        k = 0
        for k in range(steps):
            if self.optimize(tractor, **kwargs)[0] <= dchisq:
                break
        return dict(steps=k)

    def getUpdateDirection(self, tractor, allderivs, damp=0., priors=True, scale_columns=True,
                           scales_only=False, chiImages=None, variance=False,
                           shared_params=True, get_A_matrix=False):
        """The least-squares step X of tractor/lsqr_optimizer.py:328-644 (see
        the module docstring); [] when nothing constrains it, None on
        non-finite derivatives; (X, 1 / column norms^2) with `variance`;
        the column norms alone with scales_only."""
        ncols = len(allderivs)
        sysm = _SparseSystem(ncols)
        for param in allderivs:
            for _, img in param:
                sysm.image_row(img)
        for col, param in enumerate(allderivs):
            for deriv, img in param:
                sysm.add_patch(col, deriv, img, keep=not scales_only)
        b = None
        if priors:
            pd = tractor.getLogPriorDerivatives()
            if pd is not None:
                prow, pcol, pval, pb, _ = pd
                base = sysm.nrows
                for r, c, v in zip(prow, pcol, pval):
                    sysm.add(c, np.asarray(r) + base, np.asarray(v, np.float64),
                             keep=not scales_only)
                sysm.nrows += max([max(r) for r in prow if len(r)] or [-1]) + 1
                b = np.zeros(sysm.nrows, np.float32)
                b[base:] = np.hstack(pb)
        if not any(sysm.rows) and not scales_only:
            return []
        if scales_only:
            return np.sqrt(sysm.sumsq)
        var = sysm.sumsq.copy() if variance else None
        scale = np.sqrt(sysm.sumsq) if scale_columns else None
        colmap = None
        if shared_params:
            # columns of the same underlying parameter merge
            # (lsqr_optimizer.py:367-376, 495-505)
            p_now = tractor.getParams()
            tractor.setParams(np.arange(len(p_now)))
            ids = tractor.getParams()
            tractor.setParams(p_now)
            _, colmap = np.unique(ids, return_inverse=True)
        cols = {}
        for c, r, v in sysm.entries():
            if scale is not None:
                if scale[c] == 0:
                    continue
                v = v / scale[c]
                keep = np.flatnonzero(np.abs(v) > 1e-8)
                if len(keep) == 0:
                    continue
                r, v = r[keep], v[keep]
            cc = int(colmap[c]) if colmap is not None else c
            cols.setdefault(cc, []).append((r, v))
        if not cols:
            return []
        n_out = int(colmap.max()) + 1 if colmap is not None else ncols
        if b is None:
            b = np.zeros(sysm.nrows, np.float32)
        chis = dict(zip(tractor.getImages(), chiImages)) if chiImages is not None else {}
        for img, r0 in sysm.first_row.items():
            chi = chis.get(img)
            if chi is None:
                chi = tractor.getChiImage(img=img)
            chi = np.ravel(chi)
            b[r0:r0 + len(chi)] = chi
        allv = [v for lst in cols.values() for (_, v) in lst]
        if not np.all(np.isfinite(np.concatenate(allv))):
            print('Warning: infinite derivatives; bailing out')
            return None
        packed = []
        for c in range(n_out):
            lst = cols.get(c, [])
            if lst:
                packed.append((np.concatenate([r for r, _ in lst]),
                               np.concatenate([v for _, v in lst])))
            else:
                packed.append((np.zeros(0, np.int64), np.zeros(0)))
        X = solve_least_squares(packed, b.astype(np.float64), sysm.nrows, n_out, damp=damp)
        if colmap is not None:
            X = X[colmap]
        X = np.array(X, np.float64)
        if scale is not None:
            ok = scale > 0
            X[ok] /= scale[ok]
        if variance:
            return X, 1. / (var[colmap] if colmap is not None else var)
        return X
