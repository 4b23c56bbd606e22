"""Tractor: images + catalog, model images, chi, log-probability, derivatives,
optimizer dispatch (tractor/engine.py:30-497).

Model images and derivatives are planned for every source at once into one
RenderPlan (tractor_amd.plan) per call -- one device launch per render bucket
and one synchronisation, instead of one render and one copy per source --
and a line search evaluates all of its trial steps in one plan as well
(Tractor._log_probs).  Sources that have no planner (user classes) render
through their own getModelPatch / getParamDerivatives.
"""
import logging

import numpy as np

from .patch import Patch
from .plan import Pending, RenderPlan, ready, resolve
from .utils import MultiParams, _isint

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


def _chisq(img, model):
    """chi = (data - model) * inverr in the image's precision, summed as
    FP64 squares (engine.py:455-479)."""
    chi = (img.getImage() - model) * img.getInvError()
    return float((chi.astype(float) ** 2).sum())


class Tractor(MultiParams):
    """Tractor(images, catalog, optimizer=None, model_kwargs=None)."""

    @staticmethod
    def getName():
        return 'Tractor'

    @staticmethod
    def getNamedParams():
        return dict(images=0, catalog=1)

    def __init__(self, images=None, catalog=None, optimizer=None, model_kwargs=None):
        images = Images(*([] if images is None else images)) \
            if not isinstance(images, Images) else images
        catalog = Catalog(*([] if catalog is None else catalog)) \
            if not isinstance(catalog, Catalog) else catalog
        super(Tractor, self).__init__(images, catalog)
        self.modtype = np.float32
        self.modelMasks = None
        self.expectModelMasks = False
        if optimizer is None:
            from .lsqr_optimizer import LsqrOptimizer
            optimizer = LsqrOptimizer()
        self.optimizer = optimizer
        self.model_kwargs = dict(model_kwargs or {})

    def __str__(self):
        return '%s with %i sources and %i images' % (self.getName(), len(self.catalog),
                                                     len(self.images))

    def __call__(self, X):
        self.setParams(X)
        return self.getLogProb()

    # pickling keeps the model, not device state (engine.py:159-177)
    def __getstate__(self):
        return dict(version=1, images=self.images, catalog=self.catalog, liquid=self.liquid,
                    modtype=self.modtype, masks=self.modelMasks,
                    expect=self.expectModelMasks, optimizer=self.optimizer)

    def __setstate__(self, st):
        self.subs = [st['images'], st['catalog']]
        self.namedparams = dict(images=0, catalog=1)
        self.paramnames = {0: 'images', 1: 'catalog'}
        self.liquid = st['liquid']
        self.modtype = st['modtype']
        self.modelMasks = st['masks']
        self.expectModelMasks = st['expect']
        self.optimizer = st['optimizer']
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

    # ---- optimizer plugin (engine.py:210-277) ----
    def _call_optimizer(self, method, kwargs):
        kw = dict(self.model_kwargs)
        kw.update(kwargs)
        return getattr(self.optimizer, method)(self, **kw)

    def optimize(self, **kwargs):
        return self._call_optimizer('optimize', kwargs)

    def optimize_loop(self, **kwargs):
        return self._call_optimizer('optimize_loop', kwargs)

    def optimize_forced_photometry(self, **kwargs):
        return self._call_optimizer('forced_photometry', kwargs)

    # ---- model masks (engine.py:341-387) ----
    def setModelMasks(self, masks, assumeMasks=True):
        """One dict {source: ModelMask} per image, or None."""
        assert masks is None or len(masks) == len(self.images)
        self.modelMasks = masks
        self.expectModelMasks = masks is not None and assumeMasks

    def _getModelMaskFor(self, image, src):
        if self.modelMasks is None:
            return None
        return self.modelMasks[self.images.index(image)].get(src, None)

    def _mask_or_skip(self, img, src):
        """(skip, mask): with expected model masks a source without one in
        this image is not rendered there."""
        mask = self._getModelMaskFor(img, src)
        return (self.expectModelMasks and mask is None), mask

    # ---- planned renders ----
    def _plan_source_model(self, plan, img, src, minsb, kw):
        skip, mask = self._mask_or_skip(img, src)
        if skip:
            return ready(None)
        fn = getattr(src, '_plan_model', None)
        if fn is None:
            return ready(src.getModelPatch(img, modelMask=mask, minsb=minsb, **kw))
        return fn(plan, img, minsb=minsb, modelMask=mask, **kw)

    def _plan_model_image(self, plan, img, srcs, sky, minsb, kw):
        """A Pending model image: sky plus every source's patch, added in
        catalog order (engine.py:424-445)."""
        parts = [self._plan_source_model(plan, img, s, minsb, kw) for s in srcs if s is not None]

        def done():
            mod = np.zeros(img.getModelShape(), self.modtype)
            if sky:
                img.getSky().addTo(mod)
            for p in parts:
                patch = p.get()
                if patch is not None:
                    patch.addTo(mod)
            return mod
        return Pending(None, done)

    def _plan_source_derivs(self, plan, src, img, kw):
        skip, mask = self._mask_or_skip(img, src)
        if skip:
            return ready([None] * src.numberOfParams())
        fn = getattr(src, '_plan_derivs', None)
        if fn is None:
            return ready(src.getParamDerivatives(img, modelMask=mask, **kw))
        return fn(plan, img, modelMask=mask, **kw)

    # ---- models, chi, likelihood (engine.py:414-497) ----
    def getModelPatch(self, img, src, **kwargs):
        kw = dict(self.model_kwargs)
        kw.update(kwargs)
        minsb = kw.pop('minsb', None)
        return resolve(self._plan_source_model, img, src, minsb, kw)

    def getModelImage(self, img, srcs=None, sky=True, minsb=None, **kwargs):
        if _isint(img):
            img = self.getImage(img)
        kw = dict(self.model_kwargs)
        kw.update(kwargs)
        return resolve(self._plan_model_image, img, self.catalog if srcs is None else srcs, sky,
                       minsb, kw)

    def getModelImages(self, **kwargs):
        plan = RenderPlan()
        kw = dict(self.model_kwargs)
        kw.update(kwargs)
        minsb = kw.pop('minsb', None)
        sky = kw.pop('sky', True)
        srcs = kw.pop('srcs', None)
        ims = [self._plan_model_image(plan, img, self.catalog if srcs is None else srcs, sky,
                                      minsb, kw) for img in self.images]
        plan.run()
        for m in ims:
            yield m.get()

    def getChiImage(self, imgi=-1, img=None, srcs=None, minsb=0., **kwargs):
        if img is None:
            img = self.getImage(imgi)
        mod = self.getModelImage(img, srcs=srcs, minsb=minsb, **kwargs)
        return (img.getImage() - mod) * img.getInvError()

    def getChiImages(self, **kwargs):
        for img in self.images:
            yield self.getChiImage(img=img, **kwargs)

    def getLogLikelihood(self, **kwargs):
        plan = RenderPlan()
        kw = dict(self.model_kwargs)
        kw.update(kwargs)
        mods = [(img, self._plan_model_image(plan, img, self.catalog, True, 0., kw))
                for img in self.images]
        plan.run()
        return -0.5 * sum(_chisq(img, m.get()) for img, m in mods)

    def getLogProb(self, **kwargs):
        return self._log_probs([None], **kwargs)[0]

    def _log_probs(self, param_sets, **kwargs):
        """log-probabilities (engine.py:481-497: the prior, -inf when it is,
        plus the log-likelihood; NaN -> -inf) at each parameter vector of
        param_sets (None: the current parameters), all renders in one plan.
        The parameters are left as they were."""
        kw = dict(self.model_kwargs)
        kw.update(kwargs)
        p_now = self.getParams()
        plan = RenderPlan()
        todo = []
        try:
            for p in param_sets:
                if p is not None:
                    self.setParams(p)
                prior = self.getLogPrior()
                if prior == -np.inf:
                    todo.append((prior, None))
                    continue
                todo.append((prior, [(img, self._plan_model_image(plan, img, self.catalog, True,
                                                                  0., kw))
                                     for img in self.images]))
        finally:
            self.setParams(p_now)
        plan.run()
        out = []
        for prior, mods in todo:
            if mods is None:
                out.append(prior)
                continue
            lnp = prior - 0.5 * sum(_chisq(img, m.get()) for img, m in mods)
            out.append(-np.inf if np.isnan(lnp) else lnp)
        return out

    # ---- derivatives (engine.py:279-339) ----
    def _image_param_columns(self, img, kw):
        """Columns of one image's thawed parameters: the image's own
        derivatives, or a forward difference of its model image where it
        has none (False)."""
        cols = []
        mod0 = None
        for i, d in enumerate(img.getParamDerivatives(self, self.catalog)):
            if d is False:
                if mod0 is None:
                    mod0 = self.getModelImage(img, **kw)
                step = img.getStepSizes()[i]
                old = img.setParam(i, img.getParams()[i] + step)
                mod = self.getModelImage(img, **kw)
                img.setParam(i, old)
                d = Patch(0, 0, (mod - mod0) / step)
            cols.append([(d, img)])
        return cols

    def getDerivs(self, **kwargs):
        """One list per thawed parameter of (derivative Patch, image) pairs:
        image parameters first, then each thawed source's parameters over
        every image."""
        cols = []
        if not self.isParamFrozen('images'):
            for i in self.images.getThawedParamIndices():
                cols.extend(self._image_param_columns(self.images[i], kwargs))
        if not self.isParamFrozen('catalog'):
            plan = RenderPlan()
            todo = [(src, [(img, self._plan_source_derivs(plan, src, img, kwargs))
                           for img in self.images])
                    for src in self.catalog.getThawedSources()]
            plan.run()
            for src, per_img in todo:
                scols = [[] for _ in range(src.numberOfParams())]
                for img, p in per_img:
                    for k, d in enumerate(p.get()):
                        if d is not None:
                            scols[k].append((d, img))
                cols.extend(scols)
        assert len(cols) == self.numberOfParams()
        return cols
