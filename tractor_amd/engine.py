"""Tractor: images + catalog, model images, chi, log-probability, derivatives,
optimizer dispatch (tractor/engine.py:30-497)."""
import logging

import numpy as np

from .utils import MultiParams, _isint
from .patch import Patch

logger = logging.getLogger('tractor.engine')


def logverb(*args):
    if logger.isEnabledFor(logging.DEBUG):
        logger.debug(' '.join(map(str, args)))


def logmsg(*args):
    logger.info(' '.join(map(str, args)))


def isverbose():
    return logger.isEnabledFor(logging.DEBUG)


class Images(MultiParams):
    def getThawedSources(self):
        return self._getActiveSubs()


class Catalog(MultiParams):
    def __str__(self):
        return ('Catalog: %i sources, %i parameters' % (len(self), self.numberOfParams()))

    def getThawedSources(self):
        return self._getActiveSubs()

    def getFrozenSources(self):
        return self._getInactiveSubs()

    def getNamedParamName(self, j):
        return 'source%i' % j


class OptResult(object):
    pass


class Tractor(MultiParams):
    """Tractor(images, catalog, optimizer=None, model_kwargs=None)."""

    @staticmethod
    def getName():
        return 'Tractor'

    @staticmethod
    def getNamedParams():
        return dict(images=0, catalog=1)

    def __init__(self, images=None, catalog=None, optimizer=None, model_kwargs=None):
        images = [] if images is None else images
        catalog = [] if catalog is None else catalog
        if not isinstance(images, Images):
            images = Images(*images)
        if not isinstance(catalog, Catalog):
            catalog = Catalog(*catalog)
        super(Tractor, self).__init__(images, catalog)
        self.modtype = np.float32
        self.modelMasks = None
        self.expectModelMasks = False
        if optimizer is None:
            from .lsqr_optimizer import LsqrOptimizer
            optimizer = LsqrOptimizer()
        self.optimizer = optimizer
        self.model_kwargs = {} if model_kwargs is None else model_kwargs

    def __str__(self):
        return '%s with %i sources and %i images' % (self.getName(), len(self.catalog),
                                                     len(self.images))

    def __call__(self, X):
        self.setParams(X)
        return self.getLogProb()

    def __getstate__(self):
        return (1, self.getImages(), self.getCatalog(), self.liquid, self.modtype,
                self.modelMasks, self.expectModelMasks, self.optimizer)

    def __setstate__(self, state):
        (ver, images, catalog, self.liquid, self.modtype, self.modelMasks,
         self.expectModelMasks, self.optimizer) = state
        self.subs = [images, catalog]
        self.namedparams = dict(images=0, catalog=1)
        self.paramnames = {0: 'images', 1: 'catalog'}
        self.model_kwargs = {}

    def getNImages(self):
        return len(self.images)

    def getImage(self, imgi):
        return self.images[imgi]

    def getImages(self):
        return self.images

    def getCatalog(self):
        return self.catalog

    def setCatalog(self, srcs):
        self.catalog = srcs

    def setImages(self, ims):
        self.images = ims

    def addImage(self, img):
        self.images.append(img)

    def addSource(self, src):
        self.catalog.append(src)

    def addSources(self, srcs):
        self.catalog.extend(srcs)

    def removeSource(self, src):
        self.catalog.remove(src)

    # ---- optimizer dispatch (engine.py:210-277)
    def optimize_forced_photometry(self, **kwargs):
        kw = self.model_kwargs.copy()
        kw.update(kwargs)
        return self.optimizer.forced_photometry(self, **kw)

    def optimize(self, **kwargs):
        kw = self.model_kwargs.copy()
        kw.update(kwargs)
        return self.optimizer.optimize(self, **kw)

    def optimize_loop(self, **kwargs):
        kw = self.model_kwargs.copy()
        kw.update(kwargs)
        return self.optimizer.optimize_loop(self, **kw)

    # ---- derivatives (engine.py:279-339)
    def getDerivs(self, **kwargs):
        allderivs = []
        srcs = [] if self.isParamFrozen('catalog') else list(self.catalog.getThawedSources())
        allsrcs = self.catalog
        if not self.isParamFrozen('images'):
            for i in self.images.getThawedParamIndices():
                img = self.images[i]
                derivs = img.getParamDerivatives(self, allsrcs)
                mod0 = None
                for di, deriv in enumerate(derivs):
                    if deriv is False:
                        if mod0 is None:
                            mod0 = self.getModelImage(img, **kwargs)
                            p0 = img.getParams()
                            stepsizes = img.getStepSizes()
                        old = img.setParam(di, p0[di] + stepsizes[di])
                        mod = self.getModelImage(img, **kwargs)
                        img.setParam(di, old)
                        deriv = Patch(0, 0, (mod - mod0) / stepsizes[di])
                    allderivs.append([(deriv, img)])
        for src in srcs:
            srcderivs = [[] for i in range(src.numberOfParams())]
            for img in self.images:
                derivs = self._getSourceDerivatives(src, img, **kwargs)
                for k, deriv in enumerate(derivs):
                    if deriv is None:
                        continue
                    srcderivs[k].append((deriv, img))
            allderivs.extend(srcderivs)
        assert len(allderivs) == self.numberOfParams()
        return allderivs

    def setModelMasks(self, masks, assumeMasks=True):
        self.modelMasks = masks
        assert (masks is None) or (len(masks) == len(self.images))
        self.expectModelMasks = (masks is not None) and assumeMasks

    def _getModelMaskFor(self, image, src):
        if self.modelMasks is None:
            return None
        i = self.images.index(image)
        return self.modelMasks[i].get(src, None)

    def _getSourceDerivatives(self, src, img, **kwargs):
        mask = self._getModelMaskFor(img, src)
        if self.expectModelMasks and mask is None:
            return [None] * src.numberOfParams()
        return src.getParamDerivatives(img, modelMask=mask, **kwargs)

    # ---- models, chi, likelihood (engine.py:414-497)
    def getModelPatch(self, img, src, **kwargs):
        mask = self._getModelMaskFor(img, src)
        if self.expectModelMasks and mask is None:
            return None
        kw = self.model_kwargs.copy()
        kw.update(kwargs)
        return src.getModelPatch(img, modelMask=mask, **kw)

    def getModelImage(self, img, srcs=None, sky=True, minsb=None, **kwargs):
        if _isint(img):
            img = self.getImage(img)
        mod = np.zeros(img.getModelShape(), self.modtype)
        if sky:
            img.getSky().addTo(mod)
        if srcs is None:
            srcs = self.catalog
        for src in srcs:
            if src is None:
                continue
            patch = self.getModelPatch(img, src, minsb=minsb, **kwargs)
            if patch is None:
                continue
            patch.addTo(mod)
        return mod

    def getModelImages(self, **kwargs):
        for img in self.images:
            yield self.getModelImage(img, **kwargs)

    def getChiImages(self, **kwargs):
        for img in self.images:
            yield self.getChiImage(img=img, **kwargs)

    def getChiImage(self, imgi=-1, img=None, srcs=None, minsb=0., **kwargs):
        if img is None:
            img = self.getImage(imgi)
        mod = self.getModelImage(img, srcs=srcs, minsb=minsb, **kwargs)
        return (img.getImage() - mod) * img.getInvError()

    def getLogLikelihood(self, **kwargs):
        chisq = 0.
        for chi in self.getChiImages(**kwargs):
            chisq += (chi.astype(float) ** 2).sum()
        return -0.5 * chisq

    def getLogProb(self, **kwargs):
        lnprior = self.getLogPrior()
        if lnprior == -np.inf:
            return lnprior
        lnp = lnprior + self.getLogLikelihood(**kwargs)
        if np.isnan(lnp):
            return -np.inf
        return lnp
