"""Patch (an image cut-out with an origin) and ModelMask
(tractor/patch.py:3-446 semantics)."""
import numpy as np

from .render import get_overlapping_region


class ModelMask(object):
    """ModelMask(x0, y0, w, h) or ModelMask(x0, y0, mask) (tractor/patch.py:3-56)."""
    def __init__(self, x0, y0, *args):
        self.x0 = x0
        self.y0 = y0
        if len(args) == 2:
            self.w, self.h = args
            self.mask = None
        elif len(args) == 1:
            self.mask = args[0]
            self.h, self.w = self.mask.shape
        else:
            raise ValueError('Wrong number of arguments')

    def __str__(self):
        s = 'ModelMask: origin (%i,%i), w=%i, h=%i' % (self.x0, self.y0, self.w, self.h)
        return s if self.mask is None else s + ', has mask'

    __repr__ = __str__

    @staticmethod
    def fromExtent(x0, x1, y0, y1):
        return ModelMask(x0, y0, x1 - x0, y1 - y0)

    @property
    def shape(self):
        return (self.h, self.w)

    @property
    def x1(self):
        return self.x0 + self.w

    @property
    def y1(self):
        return self.y0 + self.h

    @property
    def extent(self):
        return (self.x0, self.x0 + self.w, self.y0, self.y0 + self.h)


def add_patches(pa, pb):
    p = pa
    if pb is not None:
        p = pb if p is None else p + pb
    return p


class Patch(object):
    """Image patch with origin (x0, y0) (tractor/patch.py:71-446)."""

    def __init__(self, x0, y0, patch):
        self.x0 = x0
        self.y0 = y0
        self.patch = patch
        self.name = ''

    def __str__(self):
        s = 'Patch: '
        if self.name:
            s += self.name + ' '
        s += 'origin (%i,%i) ' % (self.x0, self.y0)
        if self.patch is not None:
            H, W = self.patch.shape
            s += 'size (%i x %i)' % (W, H)
        else:
            s += '(no image)'
        return s

    __repr__ = __str__

    def setName(self, name):
        self.name = name

    def getName(self):
        return self.name

    def copy(self):
        return Patch(self.x0, self.y0, None if self.patch is None else self.patch.copy())

    @property
    def shape(self):
        return (0, 0) if self.patch is None else self.patch.shape

    @property
    def x1(self):
        return self.x0 + self.patch.shape[1]

    @property
    def y1(self):
        return self.y0 + self.patch.shape[0]

    def getImage(self):
        return self.patch

    def getX0(self):
        return self.x0

    def getY0(self):
        return self.y0

    def getOrigin(self):
        return (self.x0, self.y0)

    def getExtent(self, margin=0):
        H, W = self.shape
        return (self.x0 - margin, self.x0 + W + margin, self.y0 - margin, self.y0 + H + margin)

    def getWidth(self):
        return self.shape[1]

    def getHeight(self):
        return self.shape[0]

    def overlapsBbox(self, bbox):
        ext = self.getExtent()
        x0, x1, y0, y1 = bbox
        return (ext[0] < x1 and ext[1] > x0 and ext[2] < y1 and ext[3] > y0)

    def hasNonzeroOverlapWith(self, other):
        if self.patch is None or other.patch is None:
            return False
        return self.hasBboxOverlapWith(other)

    def hasBboxOverlapWith(self, other):
        return self.overlapsBbox(other.getExtent())

    def clipTo(self, W, H):
        """Clip to the image [0,W) x [0,H) (tractor/patch.py:241-278)."""
        if self.patch is None:
            return False
        if self.x0 >= W or self.y0 >= H:
            self.patch = None
            return False
        if self.x0 < 0:
            self.patch = self.patch[:, -self.x0:]
            self.x0 = 0
        if self.y0 < 0:
            self.patch = self.patch[-self.y0:, :]
            self.y0 = 0
        h, w = self.patch.shape
        if self.x0 + w > W:
            self.patch = self.patch[:, :W - self.x0]
        if self.y0 + h > H:
            self.patch = self.patch[:H - self.y0, :]
        if self.patch.size == 0:
            self.patch = None
            return False
        return True

    def getSlice(self, parent=None):
        if self.patch is None:
            return ([], [])
        ph, pw = self.patch.shape
        if parent is not None:
            H, W = parent.shape
            return (slice(np.clip(self.y0, 0, H), np.clip(self.y0 + ph, 0, H)),
                    slice(np.clip(self.x0, 0, W), np.clip(self.x0 + pw, 0, W)))
        return (slice(self.y0, self.y0 + ph), slice(self.x0, self.x0 + pw))

    def getPixelIndices(self, parent):
        """Flat indices into parent of the patch pixels (tractor/patch.py:331-337)."""
        if self.patch is None:
            return np.array([], np.int64)
        h, w = self.patch.shape
        H, W = parent.shape
        X, Y = np.meshgrid(np.arange(w), np.arange(h))
        return (Y.ravel() + self.y0) * W + (X.ravel() + self.x0)

    def addTo(self, img, scale=1.):
        """img[overlap] += scale * patch (tractor/patch.py:341-354)."""
        if self.patch is None:
            return
        h, w = self.patch.shape
        H, W = img.shape
        outy, iny = get_overlapping_region(self.y0, self.y0 + h - 1, 0, H - 1)
        outx, inx = get_overlapping_region(self.x0, self.x0 + w - 1, 0, W - 1)
        if outy == [] or outx == []:
            return
        p = self.patch[iny, inx]
        if scale != 1.:
            p = p * scale
        img[outy, outx] += p

    def _arith(self, other, op):
        if isinstance(other, Patch):
            if self.patch is None:
                return other.copy() if op is np.add else Patch(other.x0, other.y0,
                                                               None if other.patch is None else -other.patch)
            if other.patch is None:
                return self.copy()
            ext = self.getExtent()
            oext = other.getExtent()
            x0, x1 = min(ext[0], oext[0]), max(ext[1], oext[1])
            y0, y1 = min(ext[2], oext[2]), max(ext[3], oext[3])
            dt = np.result_type(self.patch.dtype, other.patch.dtype)
            out = np.zeros((y1 - y0, x1 - x0), dt)
            out[self.y0 - y0:self.y1 - y0, self.x0 - x0:self.x1 - x0] = self.patch
            sl = (slice(other.y0 - y0, other.y1 - y0), slice(other.x0 - x0, other.x1 - x0))
            out[sl] = op(out[sl], other.patch)
            return Patch(x0, y0, out)
        if self.patch is None:
            return Patch(self.x0, self.y0, None)
        return Patch(self.x0, self.y0, op(self.patch, other))

    def __add__(self, other):
        return self._arith(other, np.add)

    def __sub__(self, other):
        return self._arith(other, np.subtract)

    def __mul__(self, f):
        if self.patch is None:
            return Patch(self.x0, self.y0, None)
        return Patch(self.x0, self.y0, self.patch * f)

    __rmul__ = __mul__

    def __div__(self, f):
        return Patch(self.x0, self.y0, None if self.patch is None else self.patch / f)

    __truediv__ = __div__

    def __iadd__(self, other):
        r = self + other
        self.x0, self.y0, self.patch = r.x0, r.y0, r.patch
        return self

    def __imul__(self, f):
        if self.patch is not None:
            self.patch = self.patch * f
        return self

    def sum(self):
        return 0. if self.patch is None else self.patch.sum()
