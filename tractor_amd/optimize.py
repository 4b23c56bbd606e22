"""Optimizer plugin base: line search (tryUpdates) and forced photometry
(tractor/optimize.py:7-520)."""
import numpy as np

from .engine import logverb, logmsg, OptResult


ALPHAS = np.append(2. ** np.arange(-10, 1), [np.sqrt(2.), 2.])


def walk_line_search(lnp0, lnps, visit=None):
    """The acceptance rule of the line searches (tractor/optimize.py:
    448-458, constrained_optimizer.py:57-77): walk the trials in order; a
    non-finite log-probability, or one more than 1 below the best so far,
    ends the walk; the best trial that improves on lnp0 is kept.  visit(i)
    is called on each trial reached.  Returns (index or None, best lnp)."""
    best_i, best = None, lnp0
    for i, lnp in enumerate(lnps):
        if visit is not None:
            visit(i)
        if not np.isfinite(lnp) or lnp < best - 1.:
            break
        if lnp > best:
            best_i, best = i, lnp
    return best_i, best


class Optimizer(object):
    """optimize(tractor, alphas, damp, priors, scale_columns, shared_params,
    variance, just_variance) -> (dlnp, X, alpha); optimize_loop -> dict;
    forced_photometry -> OptResult."""

    def optimize(self, tractor, alphas=None, damp=0, priors=True, scale_columns=True,
                 shared_params=True, variance=False, just_variance=False):
        raise NotImplementedError

    def optimize_loop(self, tractor, **kwargs):
        raise NotImplementedError

    def tryUpdates(self, tractor, X, alphas=None):
        """The line search along X (tractor/optimize.py:425-470) -> (dlnp,
        alpha): every trial evaluated in one plan, then walked in order with
        walk_line_search; the tractor is left at the accepted trial."""
        alphas = ALPHAS if alphas is None else alphas
        p0 = np.asarray(tractor.getParams(), np.float64)
        trials = [p0 + a * np.asarray(X, np.float64) for a in alphas]
        lnps = tractor._log_probs([None] + trials)
        i, best = walk_line_search(lnps[0], lnps[1:])
        if i is None:
            tractor.setParams(p0)
            return 0, 0.
        tractor.setParams(trials[i])
        return best - lnps[0], alphas[i]

    # ---- forced photometry (tractor/optimize.py:16-222, lsqr_optimizer.py:11-251)
    def forced_photometry(self, tractor, alphas=None, damp=0, priors=False, minsb=0., mindlnp=1.,
                          rois=None, sky=False, minFlux=None, fitstats=False, fitstat_extras=None,
                          justims0=False, variance=False, skyvariance=False, shared_params=True,
                          nonneg=False, nilcounts=-1e30, wantims=True, negfluxval=None, **kwargs):
        from .calib import LinearPhotoCal
        if rois is not None or sky:
            raise NotImplementedError('forced_photometry with rois or sky fitting')
        result = OptResult()
        imgs = tractor.getImages()
        scales = []
        for img in imgs:
            assert isinstance(img.getPhotoCal(), LinearPhotoCal)
            scales.append(img.getPhotoCal().getScale())
        Nsourceparams = tractor.catalog.numberOfParams()
        srcs = list(tractor.catalog.getThawedSources())
        umodels = self._get_umodels(tractor, srcs, imgs, minsb, **kwargs)
        fsrcs = list(tractor.catalog.getFrozenSources())
        mod0 = [tractor.getModelImage(img, fsrcs, minsb=minsb, sky=True) for img in imgs]
        derivs = [[] for i in range(Nsourceparams)]
        for tim, umods, scale in zip(imgs, umodels, scales):
            for um, dd in zip(umods, derivs):
                if um is None:
                    continue
                dd.append((um * scale, tim))
        assert len(derivs) == tractor.numberOfParams()
        lnp0 = None
        chis0 = None
        damp0 = 1e-3
        damping = damp
        ims0 = imsBest = None
        while True:
            tryAgain = False
            quitNow = False
            p0 = tractor.getParams()
            if lnp0 is None:
                lnp0, chis0, ims0 = self._lnp_for_update(tractor, mod0, imgs, umodels, None, None,
                                                         p0, scales, priors, minFlux)
            if justims0:
                result.lnp0, result.chis0, result.ims0 = lnp0, chis0, ims0
                return result
            X = self.getUpdateDirection(tractor, derivs, damp=damping, priors=priors,
                                        scale_columns=False, chiImages=chis0,
                                        shared_params=shared_params)
            if X is None or len(X) == 0:
                break
            als = ALPHAS if alphas is None else alphas
            if (minFlux is None) or np.all((np.array(p0) + X) >= minFlux):
                als = [1.]
                quitNow = True
            elif damp == 0.0:
                damping = damp0
                damp0 *= 10.
                if damp0 < 1e3:
                    tryAgain = True
            lnpBest = lnp0
            alphaBest = None
            chiBest = None
            for alpha in als:
                lnp, chis, ims = self._lnp_for_update(tractor, mod0, imgs, umodels, X, alpha, p0,
                                                      scales, priors, minFlux)
                if lnp < (lnpBest - 1.):
                    break
                if not np.isfinite(lnp):
                    break
                if lnp > lnpBest:
                    alphaBest, lnpBest, chiBest, imsBest = alpha, lnp, chis, ims
            if alphaBest is not None:
                if minFlux is not None:
                    pa = [max(minFlux, p + alphaBest * d) for p, d in zip(p0, X)]
                else:
                    pa = [p + alphaBest * d for p, d in zip(p0, X)]
                tractor.catalog.setParams(pa)
                dlogprob = lnpBest - lnp0
                lnp0, chis0 = lnpBest, chiBest
            else:
                dlogprob = 0.
            if dlogprob < mindlnp and not tryAgain:
                break
            if quitNow:
                break
        result.ims0 = ims0
        result.ims1 = imsBest
        if variance or fitstats:
            ims = imsBest if imsBest is not None else ims0
            fluxes = tractor.catalog.getParams()
            stats = [self._unit_stats(img, umods, scale, fluxes, im[1])
                     for img, umods, scale, im in zip(imgs, umodels, scales, ims)]
            if variance:
                result.IV = np.zeros(Nsourceparams)
                for a, _ in stats:
                    result.IV[:len(a)] += a[:, 9]
            if fitstats:
                result.fitstats = self._get_fitstats(imgs, stats)
        return result

    def _unit_stats(self, img, umods, scale, fluxes, mod):
        """Per-source sums of Optimizer._get_fitstats (optimize.py:233-376)
        and _get_iv (378-423) for one image on the device (trh_fp_fitstats,
        trh_fp_imstats): the unit-flux patches clipped to the image and packed
        back to back, the model with its sky taken out (the chi keeps it).
        Returns ((n, 10) array, (chisq, npix) of the image)."""
        import torch
        from . import _lib
        from ._lib import check, ptr
        lib = _lib.load()
        h = _lib.handle(0)
        dev = torch.device('cuda', 0)
        H, W = img.shape
        n = len(umods)
        win = np.zeros((n, 4), np.int32)
        poff = np.zeros(n, np.int64)
        parts = []
        off = 0
        for i, um in enumerate(umods):
            if um is None or um.patch is None:
                continue
            p = np.asarray(um.patch, np.float32)
            ph, pw = p.shape
            X0, Y0 = max(um.x0, 0), max(um.y0, 0)
            X1, Y1 = min(um.x0 + pw, W), min(um.y0 + ph, H)
            if X1 <= X0 or Y1 <= Y0:
                continue
            sub = p[Y0 - um.y0:Y1 - um.y0, X0 - um.x0:X1 - um.x0]
            win[i] = (X0, Y0, X1 - X0, Y1 - Y0)
            poff[i] = off
            off += sub.size
            parts.append(sub.ravel())
        sky = float(getattr(img.getSky(), 'val', 0.))
        # counts from the catalog's fluxes (already clamped at minFlux)
        counts = np.asarray(fluxes[:n], np.float64) * scale
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        d_p = t(np.concatenate(parts) if parts else np.zeros(1, np.float32))
        d_model = t(np.asarray(mod, np.float32) - np.float32(sky))
        d_data = t(np.asarray(img.getImage(), np.float32))
        d_ie = t(np.asarray(img.getInvError(), np.float32))
        out = torch.zeros((n, 10), dtype=torch.float64, device=dev)
        im = torch.zeros(2, dtype=torch.float64, device=dev)
        d_win, d_poff, d_counts = t(win), t(poff), t(counts)
        check(lib.trh_fp_fitstats(h.value, n, ptr(d_win), ptr(d_poff), ptr(d_p), ptr(d_counts),
                                  ptr(d_model), ptr(d_data), ptr(d_ie), W, float(scale), sky,
                                  ptr(out), None), 'fp_fitstats')
        check(lib.trh_fp_imstats(h.value, H * W, ptr(d_model), ptr(d_data), ptr(d_ie), sky,
                                 ptr(im), None), 'fp_imstats')
        return out.cpu().numpy(), im.cpu().numpy()

    def _get_fitstats(self, imgs, stats):
        """FitStats from the per-image sums (optimize.py:233-376): sums add
        over images; profracflux = sum(ratio) / max(1, sum(flux fraction
        terms)), fracin = sum(window profile) / max(1, images counted)."""
        from .forced import FitStats
        fs = FitStats()
        a = sum(st for st, _ in stats)
        fs.imchisq = np.array([im[0] for _, im in stats])
        fs.imnpix = np.array([int(im[1]) for _, im in stats])
        fs.sky = np.array([float(getattr(img.getSky(), 'val', 0.)) for img in imgs])
        fs.prochi2, fs.pronpix, fs.promasked = a[:, 0], a[:, 1], a[:, 2]
        fs.profracflux = a[:, 3] / np.maximum(1., a[:, 4])
        fs.proflux, fs.npix = a[:, 5], a[:, 6].astype(int)
        fs.fracin = a[:, 7] / np.maximum(1., a[:, 8])
        return fs

    def _get_umodels(self, tractor, srcs, imgs, minsb, **kwargs):
        """Unit-flux patches per (image, source brightness) (optimize.py:
        147-222), every image's and source's renders planned together."""
        from .plan import RenderPlan, ready
        plan = RenderPlan()
        todo = []
        for img in imgs:
            pcal = img.getPhotoCal()
            row = []
            for src in srcs:
                counts = sum([pcal.brightnessToCounts(b) for b in src.getBrightnesses()])
                mv = 1e-3 if counts <= 0 else minsb / counts
                mask = tractor._getModelMaskFor(img, src)
                fn = getattr(src, '_plan_unit_patches', None)
                if fn is None:
                    row.append([ready(u) for u in src.getUnitFluxModelPatches(
                        img, minval=mv, modelMask=mask, **kwargs)])
                else:
                    row.append(fn(plan, img, minval=mv, modelMask=mask, **kwargs))
            todo.append(row)
        plan.run()
        umodels = []
        for row in todo:
            umods = []
            for pend in row:
                ums = [p.get() for p in pend]
                for um in ums:
                    if um is not None and um.patch is not None:
                        assert np.all(np.isfinite(um.patch))
                umods.extend(ums)
            umodels.append(umods)
        return umodels

    def _getims(self, fluxes, imgs, umodels, mod0, scales, minFlux):
        """Model = mod0 + sum flux*scale*umodel; chi (optimize.py:472-520)."""
        ims = []
        for img, umods, m0, scale in zip(imgs, umodels, mod0, scales):
            mod = m0.copy()
            for f, um in zip(fluxes, umods):
                if um is None or um.patch is None:
                    continue
                if minFlux is not None:
                    f = max(f, minFlux)
                counts = f * scale
                if counts == 0.:
                    continue
                (um * counts).addTo(mod)
            ie = img.getInvError()
            im = img.getImage()
            chi = (im - mod) * ie
            ims.append((im, mod, ie, chi, None))
        return ims

    def _lnp_for_update(self, tractor, mod0, imgs, umodels, X, alpha, p0, scales, priors,
                        minFlux):
        pa = p0 if X is None else [p + alpha * d for p, d in zip(p0, X)]
        lnp = 0.
        if priors:
            lnp += tractor.getLogPrior()
            if not np.isfinite(lnp):
                return lnp, None, None
        ims = self._getims(pa, imgs, umodels, mod0, scales, minFlux)
        chisq = 0.
        chis = []
        for _, _, _, chi, _ in ims:
            chis.append(chi)
            chisq += (chi.astype(np.float64) ** 2).sum()
        return lnp - 0.5 * chisq, chis, ims

