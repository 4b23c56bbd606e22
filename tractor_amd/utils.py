"""Parameter system: flat, freeze/thaw-aware parameter vectors over nested
model objects (the reference's Params duck type, tractor/utils.py:195-1270).

Semantics kept from the reference (names, return conventions):
* getParams / setParams / numberOfParams see only thawed ("liquid") params;
* getAllParams / setAllParams ignore the freeze state;
* setParam(i, v) returns the old value;
* getStepSizes / getLowerBounds / getUpperBounds / getMaxStep are per thawed
  parameter (None where unbounded);
* Gaussian priors enter the least-squares system as extra rows
  (getLogPriorDerivatives -> rows, cols, vals, b, mu).
"""
import numpy as np


def getClassName(obj):
    name = getattr(obj.__class__, 'classname', None)
    return name if name is not None else obj.__class__.__name__


def listmax(X, default=0):
    mx = [np.max(x) for x in X if len(x)]
    return np.max(mx) if len(mx) else default


class BaseParams(object):
    """A parameter container with no parameters (tractor/utils.py:195-305)."""

    def __repr__(self):
        return getClassName(self) + repr(self.getParams())

    def __str__(self):
        return getClassName(self) + ': ' + str(self.getParams())

    def copy(self):
        return self.__class__(*self.getAllParams())

    def hashkey(self):
        return (getClassName(self),) + tuple(self.getAllParams())

    def getParamNames(self):
        return []

    def numberOfParams(self):
        return len(self.getParams())

    def getParams(self):
        return []

    def getAllParams(self):
        return self.getParams()

    def getAllStepSizes(self, *args, **kwargs):
        return self.getStepSizes(*args, **kwargs)

    def getStepSizes(self, *args, **kwargs):
        return [1.] * self.numberOfParams()

    def setAllStepSizes(self, ss):
        self.setStepSizes(ss)

    def setStepSizes(self, ss):
        pass

    def setParams(self, p):
        assert len(p) == self.numberOfParams()
        for i, pi in enumerate(p):
            self.setParam(i, pi)

    def setAllParams(self, p):
        return self.setParams(p)

    def setParam(self, i, p):
        return None

    def getLogPrior(self):
        return 0.

    def getLogPriorDerivatives(self):
        return None

    def getLowerBounds(self):
        return [None] * self.numberOfParams()

    def getUpperBounds(self):
        return [None] * self.numberOfParams()

    def getMaxStep(self):
        return [None] * self.numberOfParams()

    def getGaussianPriors(self):
        return []


class ScalarParam(BaseParams):
    """One float parameter (tractor/utils.py:307-379)."""
    stepsize = 1.
    strformat = '%g'

    def __init__(self, val=0):
        self.val = val
        self.lower = None
        self.upper = None
        self.maxstep = None

    def __str__(self):
        return getClassName(self) + ': ' + self.strformat % self.val

    def __repr__(self):
        return getClassName(self) + '(' + repr(self.val) + ')'

    def __eq__(self, other):
        return self.getValue() == other.getValue()

    def __lt__(self, other):
        return self.getValue() < other.getValue()

    __hash__ = object.__hash__

    def copy(self):
        return self.__class__(self.val)

    def getParamNames(self):
        return [getClassName(self)]

    def numberOfParams(self):
        return 1

    def getStepSizes(self, *args, **kwargs):
        return [self.stepsize]

    def setStepSizes(self, ss):
        self.stepsize = ss[0]

    def getParams(self):
        return [self.val]

    def getAllParams(self):
        return [self.val]

    def setParams(self, p):
        assert len(p) == 1
        self._set(p[0])

    def setAllParams(self, p):
        self.setParams(p)

    def setParam(self, i, p):
        assert i == 0
        old = self.val
        self._set(p)
        return old

    def _set(self, val):
        self.val = val

    def getValue(self):
        return self.val

    def setValue(self, v):
        self._set(v)

    def getLowerBounds(self):
        return [self.lower]

    def getUpperBounds(self):
        return [self.upper]

    def getMaxStep(self):
        return [self.maxstep]


class NamedParams(object):
    """Mix-in: names for parameter slots and a per-slot thawed ("liquid")
    flag (tractor/utils.py:389-508)."""

    @staticmethod
    def getNamedParams():
        return {}

    def __new__(cls, *args, **kwargs):
        self = super(NamedParams, cls).__new__(cls)
        self.namedparams = {}
        self.paramnames = {}
        self.addNamedParams(**self.getNamedParams())
        return self

    def __init__(self):
        super(NamedParams, self).__init__()
        self.liquid = [True] * self._numberOfThings()

    def _addNamedParams(self, alias, **d):
        self.namedparams.update(d)
        if not alias:
            self.paramnames.update(dict((v, k) for k, v in d.items()))
        for n in d:
            if hasattr(self.__class__, n):
                continue
            setattr(self.__class__, n, property(
                (lambda nm: lambda x: x._getThing(x.namedparams[nm]))(n),
                (lambda nm: lambda x, v: x._setThing(x.namedparams[nm], v))(n),
                None, 'named param %s' % n))

    def addParamAliases(self, **d):
        self._addNamedParams(True, **d)

    def addNamedParams(self, **d):
        self._addNamedParams(False, **d)

    def getNamedParamIndex(self, name):
        return self.namedparams.get(name, None)

    def getNamedParamName(self, ii):
        return self.paramnames.get(ii, None)

    def _iterNamesAndVals(self):
        for i, val in enumerate(self._getThings()):
            yield (self.paramnames.get(i, None), val)

    def _index(self, paramname):
        if isinstance(paramname, (int, np.integer)):
            return int(paramname)
        i = self.getNamedParamIndex(paramname)
        if i is None:
            raise KeyError('no parameter named "%s" in %s' % (paramname, getClassName(self)))
        return i

    # freeze / thaw
    def freezeParam(self, paramname):
        self.liquid[self._index(paramname)] = False

    def freezeParams(self, *args):
        for a in args:
            self.freezeParam(a)

    def thawParam(self, paramname):
        self.liquid[self._index(paramname)] = True

    def thawParams(self, *args):
        for a in args:
            self.thawParam(a)

    def freezeAllBut(self, *args):
        self.freezeAllParams()
        self.thawParams(*args)

    def freezeAllParams(self):
        self.liquid = [False] * len(self.liquid)

    def thawAllParams(self):
        self.liquid = [True] * len(self.liquid)

    def freezeParamsRecursive(self, *pnames):
        for name, sub in self._iterNamesAndVals():
            if hasattr(sub, 'freezeParamsRecursive'):
                sub.freezeParamsRecursive(*pnames)
            if name in pnames:
                self.freezeParam(name)
        if '*' in pnames:
            self.freezeAllParams()

    def thawParamsRecursive(self, *pnames):
        for name, sub in self._iterNamesAndVals():
            if hasattr(sub, 'thawParamsRecursive'):
                sub.thawParamsRecursive(*pnames)
            if name in pnames:
                self.thawParam(name)
        if '*' in pnames:
            self.thawAllParams()

    def thawAllRecursive(self):
        self.thawParamsRecursive('*')

    def freezeAllRecursive(self):
        self.freezeParamsRecursive('*')

    def isParamFrozen(self, paramname):
        return not self.liquid[self._index(paramname)]

    def isParamThawed(self, paramname):
        return self.liquid[self._index(paramname)]

    def getFrozenParams(self):
        return [self.getNamedParamName(i) for i in self.getFrozenParamIndices()]

    def getThawedParams(self):
        return [self.getNamedParamName(i) for i in self.getThawedParamIndices()]

    def getFrozenParamIndices(self):
        return [i for i, l in enumerate(self.liquid) if not l]

    def getThawedParamIndices(self):
        return [i for i, l in enumerate(self.liquid) if l]

    def getThawedParamIndicesAndNames(self):
        return [(i, self.getNamedParamName(j)) for i, j in self._indexBoth()]

    def getLiquidIndexOfIndex(self, i):
        if not self.liquid[i]:
            return -1
        return sum(1 for l in self.liquid[:i] if l)

    def getLiquidIndex(self, paramname):
        return self.getLiquidIndexOfIndex(self._index(paramname))

    def _getLiquidArray(self, array):
        return [a for a, l in zip(array, self.liquid) if l]

    def _getFrozenArray(self, array):
        return [a for a, l in zip(array, self.liquid) if not l]

    def _enumerateLiquidArray(self, array):
        return [(i, a) for i, (a, l) in enumerate(zip(array, self.liquid)) if l]

    def _countLiquid(self):
        return sum(1 for l in self.liquid if l)

    def _indexLiquid(self, j):
        """Raw index of the j-th thawed slot."""
        k = 0
        for i, l in enumerate(self.liquid):
            if l:
                if k == j:
                    return i
                k += 1
        raise IndexError(j)

    def _indexBoth(self):
        i = 0
        for j, v in enumerate(self.liquid):
            if v:
                yield (i, j)
                i += 1


class _GaussianPriors(object):
    """Gaussian priors on named parameters (tractor/utils.py:65-133)."""
    def __init__(self, param):
        self.terms = []
        self.param = param

    def add(self, name, mu, sigma, param=None):
        if param is None:
            param = self.param
        i = param.getNamedParamIndex(name)
        if i is None:
            raise KeyError('GaussianPriors.add: parameter not found: "%s"' % name)
        self.terms.append((name, i, mu, sigma))

    def getLogPrior(self, param=None):
        param = self.param if param is None else param
        p = param.getAllParams()
        return -0.5 * sum((p[i] - mu) ** 2 / sigma ** 2 for _, i, mu, sigma in self.terms)

    def getGaussianPriors(self):
        return [(i, mu, sigma) for _, i, mu, sigma in self.terms]

    def getDerivs(self, param=None):
        param = self.param if param is None else param
        rows, cols, vals, bs, mus = [], [], [], [], []
        row0 = 0
        p = param.getParams()
        for _, j, mu, sigma in self.terms:
            i = param.getLiquidIndexOfIndex(j)
            if i == -1:
                continue
            cols.append(i)
            vals.append([1. / sigma])
            rows.append([row0])
            bs.append([-(p[i] - mu) / sigma])
            mus.append([mu])
            row0 += 1
        return rows, cols, vals, bs, mus


class GaussianPriorsMixin(object):
    def __init__(self, *args, **kwargs):
        super(GaussianPriorsMixin, self).__init__(*args, **kwargs)
        self.gpriors = _GaussianPriors(self)

    def addGaussianPrior(self, name, mu, sigma):
        self.gpriors.add(name, mu, sigma, param=self)

    def getLogPriorDerivatives(self):
        r = self.gpriors.getDerivs(param=self)
        return None if len(r[0]) == 0 else r

    def getGaussianPriors(self):
        return self.gpriors.getGaussianPriors()

    def isLegal(self):
        return True

    def getLogPrior(self):
        if not self.isLegal():
            return -np.inf
        return self.gpriors.getLogPrior(param=self)


class ParamList(GaussianPriorsMixin, NamedParams, BaseParams):
    """Values held in a list (tractor/utils.py:710-836)."""

    def __init__(self, *args):
        self.vals = list(args)
        self.lowers = [None] * len(self.vals)
        self.uppers = [None] * len(self.vals)
        self.maxstep = [None] * len(self.vals)
        super(ParamList, self).__init__()

    def copy(self):
        c = self.__class__(*self.vals)
        c.liquid = list(self.liquid)
        return c

    def hashkey(self):
        return (getClassName(self),) + tuple(self.vals)

    def getFormatString(self, i):
        return '%g'

    def __str__(self):
        ss = []
        for i, (name, val) in enumerate(self._iterNamesAndVals()):
            fmt = self.getFormatString(i)
            ss.append(('%s=' + fmt) % (name, val) if name is not None else fmt % val)
        return getClassName(self) + ': ' + ', '.join(ss)

    def __repr__(self):
        return getClassName(self) + repr(self.vals)

    def getParamNames(self):
        out = []
        for i, j in self._indexBoth():
            nm = self.getNamedParamName(j)
            out.append(nm if nm is not None else 'param%i' % i)
        return out

    def _setThing(self, i, val):
        self.vals[i] = val

    def _getThing(self, i):
        return self.vals[i]

    def _getThings(self):
        return self.vals

    def _numberOfThings(self):
        return len(self.vals)

    def setParam(self, i, val):
        ii = self._indexLiquid(i)
        old = self.vals[ii]
        self.vals[ii] = val
        return old

    def setParams(self, p):
        for i, j in self._indexBoth():
            self.vals[j] = p[i]

    def numberOfParams(self):
        return self._countLiquid()

    def getParams(self):
        return list(self._getLiquidArray(self.vals))

    def getAllParams(self):
        return list(self.vals)

    def getParam(self, i):
        return self.vals[self._indexLiquid(i)]

    def setAllParams(self, p):
        for i, pp in enumerate(p):
            self.vals[i] = pp

    def getStepSizes(self, *args, **kwargs):
        ss = getattr(self, 'stepsizes', None)
        if ss is None:
            ss = self.getAllStepSizes(*args, **kwargs)
        return list(self._getLiquidArray(ss))

    def getAllStepSizes(self, *args, **kwargs):
        ss = getattr(self, 'stepsizes', None)
        return ss if ss is not None else [1.] * len(self.vals)

    def setStepSizes(self, ss):
        if getattr(self, 'stepsizes', None) is None:
            self.stepsizes = [1.] * len(self.vals)
        for i, j in self._indexBoth():
            self.stepsizes[j] = ss[i]

    def setAllStepSizes(self, ss):
        self.stepsizes = list(ss)

    def getLowerBounds(self):
        return list(self._getLiquidArray(self.lowers))

    def getUpperBounds(self):
        return list(self._getLiquidArray(self.uppers))

    def getMaxStep(self):
        return list(self._getLiquidArray(self.maxstep))

    def __len__(self):
        return self.numberOfParams()

    def __getitem__(self, i):
        return self.getParam(i)

    def __iter__(self):
        return iter(self.getParams())


class MultiParams(BaseParams, NamedParams):
    """Params made of sub-Params (tractor/utils.py:923-1270)."""

    def __init__(self, *args):
        self.subs = list(args)
        super(MultiParams, self).__init__()

    def copy(self):
        x = self.__class__(*[s.copy() if s is not None else None for s in self.subs])
        x.liquid = list(self.liquid)
        return x

    def append(self, x):
        self.subs.append(x)
        self.liquid.append(True)

    def prepend(self, x):
        self.subs = [x] + self.subs
        self.liquid = [True] + self.liquid

    def extend(self, x):
        self.subs.extend(x)
        self.liquid.extend([True] * len(x))

    def remove(self, x):
        i = self.subs.index(x)
        del self.subs[i]
        del self.liquid[i]

    def index(self, x):
        return self.subs.index(x)

    def __len__(self):
        return len(self.subs)

    def __getitem__(self, key):
        return self.subs[key]

    def __setitem__(self, key, val):
        self.subs[key] = val

    def __iter__(self):
        return iter(self.subs)

    def hashkey(self):
        return (getClassName(self),) + tuple(None if s is None else s.hashkey() for s in self.subs)

    def __str__(self):
        s = []
        for n, v in self._iterNamesAndVals():
            s.append(str(v) if n is None else '%s: %s' % (n, str(v)))
        return getClassName(self) + ': ' + ', '.join(s)

    def _setThing(self, i, val):
        self.subs[i] = val

    def _getThing(self, i):
        return self.subs[i]

    def _getThings(self):
        return self.subs

    def _numberOfThings(self):
        return len(self.subs)

    def _getActiveSubs(self):
        return [s for s in self._getLiquidArray(self.subs) if s is not None]

    def _getInactiveSubs(self):
        return [s for s in self._getFrozenArray(self.subs) if s is not None]

    def getParamNames(self):
        n = []
        for i, s in self._enumerateLiquidArray(self.subs):
            if s is None:
                continue
            pre = self.getNamedParamName(i)
            pre = pre if pre is not None else 'param%i' % i
            n.extend('%s.%s' % (pre, post) for post in s.getParamNames())
        return n

    def printThawedParams(self):
        for nm, val in zip(self.getParamNames(), self.getParams()):
            print('  ', nm, '=', val)

    def numberOfParams(self):
        return sum(s.numberOfParams() for s in self._getActiveSubs())

    def getParams(self):
        p = []
        for s in self._getActiveSubs():
            p.extend(s.getParams())
        return p

    def getAllParams(self):
        p = []
        for s in self.subs:
            if s is not None:
                p.extend(s.getAllParams())
        return p

    def setAllParams(self, p):
        i = 0
        for s in self.subs:
            if s is None:
                continue
            n = len(s.getAllParams())
            s.setAllParams(p[i:i + n])
            i += n

    def setParams(self, p):
        i = 0
        for s in self._getActiveSubs():
            n = s.numberOfParams()
            s.setParams(p[i:i + n])
            i += n

    def setParam(self, i, p):
        off = 0
        for s in self._getActiveSubs():
            n = s.numberOfParams()
            if i < off + n:
                return s.setParam(i - off, p)
            off += n
        raise RuntimeError('setParam(%i) on %s with %i params' % (i, getClassName(self),
                                                                  self.numberOfParams()))

    def _concat(self, meth, *args, **kwargs):
        p = []
        for s in self._getActiveSubs():
            p.extend(getattr(s, meth)(*args, **kwargs))
        return p

    def getLowerBounds(self):
        return self._concat('getLowerBounds')

    def getUpperBounds(self):
        return self._concat('getUpperBounds')

    def getMaxStep(self):
        return self._concat('getMaxStep')

    def getStepSizes(self, *args, **kwargs):
        return self._concat('getStepSizes', *args, **kwargs)

    def getAllStepSizes(self, *args, **kwargs):
        p = []
        for s in self.subs:
            if s is not None:
                p.extend(s.getAllStepSizes(*args, **kwargs))
        return p

    def setStepSizes(self, ss):
        off = 0
        for s in self._getActiveSubs():
            n = s.numberOfParams()
            s.setStepSizes(ss[off:off + n])
            off += n

    def getGaussianPriors(self):
        p, off = [], 0
        for s in self._getActiveSubs():
            p.extend([(off + i, mu, sig) for i, mu, sig in s.getGaussianPriors()])
            off += s.numberOfParams()
        return p

    def getLogPrior(self):
        return sum(s.getLogPrior() for s in self._getActiveSubs())

    def getLogPriorDerivatives(self):
        rA, cA, vA, pb, mub = [], [], [], [], []
        r0 = c0 = 0
        for s in self._getActiveSubs():
            X = s.getLogPriorDerivatives()
            if X is None:
                c0 += s.numberOfParams()
                continue
            r, c, v, b, m = X
            rA.extend([[rij + r0 for rij in ri] for ri in r])
            cA.extend([ci + c0 for ci in c])
            vA.extend(v)
            pb.extend(b)
            mub.extend(m)
            c0 += s.numberOfParams()
            r0 += listmax(r, -1) + 1
        if rA == []:
            return None
        return rA, cA, vA, pb, mub


def _isint(i):
    return isinstance(i, (int, np.integer))
