"""Deferred device renders for the object API.

A RenderPlan collects every render a model image, a set of derivatives or a
whole line search needs -- FFT-path galaxy renders, pixelized-PSF point
renders, pixel-space mixture evaluations, Lanczos-shifted stamps -- and then
makes one library call per render bucket (FFT size and PSF; PSF stamp;
mixtures; image shape) into a single device buffer, with one
synchronisation and one copy back.

Each request returns a Pending.  Its output window (`extent`, x0, y0, w, h)
is fixed when the request is planned -- the reference's extent rules decide
every window before any pixel is computed -- so a caller can plan renders
that depend on another render's window (the shape-derivative renders on the
first render's extent) in the same batch.  The pixels are read with
Pending.get() once the plan has run.

This is the object API's counterpart of the reference's per-source SWIG
calls (mp_fourier.gaussian_fourier_transform_zero_mean + irfftn +
lanczos_shift_3f, tractor/galaxy.py:336-508; mix.c_gauss_2d_grid /
c_gauss_2d_masked, tractor/mixture_profiles.py:270-343;
lanczos_shift_image, tractor/psf.py:26-58): the same renders, issued per
batch instead of per source.
"""
import numpy as np
import torch

from . import _lib, render as R
from ._lib import FFT_JOB, MOG_JOB, TRH_STORE, check, ptr


class Pending(object):
    """A render result available after its plan has run.  extent: (x0, y0,
    w, h) of the result, or None when the result is None."""
    __slots__ = ('extent', 'has_derivs', '_fn', '_val', '_done')

    def __init__(self, extent, fn, has_derivs=False):
        self.extent = extent
        self.has_derivs = has_derivs
        self._fn = fn
        self._val = None
        self._done = False

    def get(self):
        if not self._done:
            self._val = self._fn()
            self._fn = None
            self._done = True
        return self._val


def ready(value, extent=None, has_derivs=False):
    """A Pending whose value is already known."""
    p = Pending(extent, None, has_derivs)
    p._val = value
    p._done = True
    return p


def _spectra_of(psf):
    """Device stamp and spectra of a pixelized PSF, cached on the PSF object
    (PixelizedPSF.fftcache, tractor/psf.py:268-280)."""
    sp = getattr(psf, '_trh_spectra', None)
    if sp is None:
        sp = R.PsfSpectra(psf.img, 'cuda')
        psf._trh_spectra = sp
    return sp


class RenderPlan(object):
    """Render requests for one batch of device calls (see the module
    docstring).  Requests return slot numbers; Pending objects built on them
    read plan.result(slot) after run()."""

    # process-wide counts of plans run with at least one request and of
    # library render calls (tests read them to check the batching)
    stats = dict(runs=0, calls=0)

    def __init__(self, device=0):
        self.device = device
        self._slots = []          # per slot: (w, h) of its output
        self._fft = {}            # (id(spectra), sz) -> (spectra, [(job, slot)])
        self._point = {}          # id(spectra) -> (spectra, [(job, slot)])
        self._mog = []            # (slot, x0, y0, w, h, amp, mean, var, fx, fy, mask, derivs, masked)
        self._shift = {}          # (H, W) -> [(img, dx, dy, slot)]
        self._host = None
        self._dx = self._dy = None
        self._ran = False
        self.launches = 0

    # ---- requests ----
    def _slot(self, w, h):
        if self._ran:
            raise RuntimeError('RenderPlan already ran')
        self._slots.append((int(w), int(h)))
        return len(self._slots) - 1

    def fft(self, psf, sz, job, w, h):
        """An FFT-path render job (trh_fft_job, out fields filled at run)
        of the pixelized PSF `psf` at FFT size sz into a w x h window."""
        sp = _spectra_of(psf)
        s = self._slot(w, h)
        if w > 0 and h > 0:
            self._fft.setdefault((id(sp), sz), (sp, []))[1].append((job, s))
        return s

    def point(self, psf, job, w, h):
        """A Lanczos-shifted stamp of the pixelized PSF `psf` into a w x h
        window (trh_render_point_batch job)."""
        sp = _spectra_of(psf)
        s = self._slot(w, h)
        if w > 0 and h > 0:
            self._point.setdefault(id(sp), (sp, []))[1].append((job, s))
        return s

    def mog(self, amp, mean, var, x0, y0, w, h, fx=0., fy=0., mask=None, derivs=False,
            masked_rules=False):
        """A pixel-space mixture on the window [x0, x0+w) x [y0, y0+h),
        components at mean + (fx, fy): c_gauss_2d_grid rules, or with
        masked_rules c_gauss_2d_masked (mask, derivatives)."""
        s = self._slot(w, h)
        if w <= 0 or h <= 0:
            return s
        self._mog.append((s, int(x0), int(y0), int(w), int(h), np.asarray(amp, np.float64),
                          np.asarray(mean, np.float64).reshape(-1, 2),
                          np.asarray(var, np.float64).reshape(-1, 2, 2), float(fx), float(fy),
                          mask, bool(derivs), bool(masked_rules)))
        return s

    def shift(self, img, dx, dy):
        """lanczos_shift_image of a float32 image (tractor/psf.py:26-58)."""
        img = np.ascontiguousarray(img, np.float32)
        H, W = img.shape
        s = self._slot(W, H)
        self._shift.setdefault((H, W), []).append((img, float(dx), float(dy), s))
        return s

    # ---- execution ----
    def run(self):
        """Issue every request (one library call per bucket), synchronise
        once and copy the results back."""
        if self._ran:
            return self
        self._ran = True
        n = len(self._slots)
        if n == 0:
            return self
        lib = _lib.load()
        h = _lib.handle(self.device).value
        dev = torch.device('cuda', self.device)
        sizes = np.array([w * hh for (w, hh) in self._slots], np.int64)
        off = np.zeros(n, np.int64)
        # slots in the order of their buckets: each shift group contiguous
        order = []
        for grp in self._shift.values():
            order.extend(s for (_, _, _, s) in grp)
        in_shift = set(order)
        order.extend(s for s in range(n) if s not in in_shift)
        pos = 0
        for s in order:
            off[s] = pos
            pos += sizes[s]
        nmog = sum(1 for m in self._mog if m[11])
        # MoG derivative planes mirror the whole buffer (same indices)
        total = max(int(pos), 1)
        buf = torch.zeros(total * (3 if nmog else 1), dtype=torch.float32, device=dev)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        keep = []

        # FFT renders, grouped by (spectra, size)
        groups = list(self._fft.values())
        if groups:
            recs = []
            starts = []
            for sp, jobs in groups:
                starts.append(sum(len(r) for r in recs))
                arr = np.zeros(len(jobs), FFT_JOB)
                for i, (J, s) in enumerate(jobs):
                    arr[i] = J
                    arr[i]['out_offset'] = off[s]
                    arr[i]['out_stride'] = self._slots[s][0]
                    arr[i]['mode'] = TRH_STORE
                    arr[i]['psf'] = 0
                recs.append(arr)
            dj = t(np.concatenate(recs).view(np.uint8))
            keep.append(dj)
            for (key, (sp, jobs)), st in zip(self._fft.items(), starts):
                sz = key[1]
                spec, _, _ = sp.get(sz)
                check(lib.trh_render_fft_batch(h, sz, ptr(dj) + FFT_JOB.itemsize * st, len(jobs),
                                               ptr(spec), ptr(buf), None, None, None, None),
                      'trh_render_fft_batch')
                self.launches += 1

        # point renders, grouped by stamp
        if self._point:
            recs = []
            starts = []
            for sp, jobs in self._point.values():
                starts.append(sum(len(r) for r in recs))
                arr = np.zeros(len(jobs), FFT_JOB)
                for i, (J, s) in enumerate(jobs):
                    arr[i] = J
                    arr[i]['out_offset'] = off[s]
                    arr[i]['out_stride'] = self._slots[s][0]
                    arr[i]['mode'] = TRH_STORE
                    arr[i]['psf'] = 0
                recs.append(arr)
            dj = t(np.concatenate(recs).view(np.uint8))
            keep.append(dj)
            for (sp, jobs), st in zip(self._point.values(), starts):
                ph, pw = sp.shape
                check(lib.trh_render_point_batch(h, ptr(dj) + FFT_JOB.itemsize * st, len(jobs),
                                                 ptr(sp.stamp), ph, pw, ptr(buf), None, None,
                                                 None, None), 'trh_render_point_batch')
                self.launches += 1

        # pixel-space mixtures: every job of the plan in one component table
        if self._mog:
            self._run_mog(lib, h, dev, buf, off, total, keep)

        # Lanczos shifts, one call per image shape
        for (H, W), grp in self._shift.items():
            imgs = np.stack([g[0] for g in grp])
            din = t(imgs.ravel())
            ddx = t(np.array([g[1] for g in grp], np.float64))
            ddy = t(np.array([g[2] for g in grp], np.float64))
            o = int(off[grp[0][3]])
            keep.extend([din, ddx, ddy])
            check(lib.trh_lanczos_shift_batch(h, ptr(din), ptr(buf) + 4 * o, len(grp), H, W,
                                              ptr(ddx), ptr(ddy), None), 'trh_lanczos_shift_batch')
            self.launches += 1

        RenderPlan.stats['runs'] += 1
        RenderPlan.stats['calls'] += self.launches
        host = buf.cpu().numpy()      # the one synchronisation
        self._host = host[:total]
        if nmog:
            self._dx = host[total:2 * total]
            self._dy = host[2 * total:]
        self._off = off
        return self

    def _run_mog(self, lib, h, dev, buf, off, total, keep):
        jobs = np.zeros(len(self._mog), MOG_JOB)
        amps, cxi, cyi, cxf, cyf, v3 = [], [], [], [], [], []
        masks = []
        mask_pos = 0
        comp = 0
        maxpix = 1
        for i, (s, x0, y0, w, hh, amp, mean, var, fx, fy, mask, derivs, masked) in \
                enumerate(self._mog):
            K = len(amp)
            cx = mean[:, 0] + fx
            cy = mean[:, 1] + fy
            ix = np.floor(cx).astype(np.int32)
            iy = np.floor(cy).astype(np.int32)
            amps.append(amp.astype(np.float32))
            cxi.append(ix)
            cyi.append(iy)
            cxf.append((cx - ix).astype(np.float32))
            cyf.append((cy - iy).astype(np.float32))
            v3.append(np.stack([var[:, 0, 0], 0.5 * (var[:, 0, 1] + var[:, 1, 0]),
                                var[:, 1, 1]], 1).astype(np.float32))
            J = jobs[i]
            J['out_offset'] = off[s]
            J['out_stride'] = w
            J['x0'], J['y0'], J['w'], J['h'] = x0, y0, w, hh
            J['comp0'], J['K'] = comp, K
            J['cutoff_mode'] = 1 if masked else 0
            J['mode'] = TRH_STORE
            J['want_derivs'] = 1 if derivs else 0
            J['scale'] = 1.
            if mask is not None:
                # the kernel reads mask[mask_offset + oy * out_stride + ox]
                m = np.ascontiguousarray(mask, np.uint8).ravel()
                assert len(m) == w * hh
                J['mask_offset'] = mask_pos
                masks.append(m)
                mask_pos += len(m)
            else:
                J['mask_offset'] = -1
            comp += K
            maxpix = max(maxpix, w * hh)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        d = [t(np.concatenate(a)) for a in (amps, cxi, cyi, cxf, cyf, v3)]
        d_mask = t(np.concatenate(masks)) if masks else None
        keep.extend(d)
        keep.append(d_mask)
        nder = any(m[11] for m in self._mog)
        dx = ptr(buf) + 4 * total if nder else None
        dy = ptr(buf) + 8 * total if nder else None
        for c0 in range(0, len(jobs), 65535):
            dj = t(jobs[c0:c0 + 65535].view(np.uint8))
            keep.append(dj)
            check(lib.trh_render_mog_batch(
                h, ptr(dj), min(65535, len(jobs) - c0), maxpix, ptr(d[0]), ptr(d[1]), ptr(d[2]),
                ptr(d[3]), ptr(d[4]), ptr(d[5]), ptr(d_mask), ptr(buf), dx, dy, None),
                'trh_render_mog_batch')
            self.launches += 1

    # ---- results ----
    def result(self, slot):
        """float32 (h, w) pixels of a slot (after run)."""
        w, hh = self._slots[slot]
        if w * hh == 0:
            return np.zeros((max(hh, 0), max(w, 0)), np.float32)
        o = int(self._off[slot])
        return self._host[o:o + w * hh].reshape(hh, w).copy()

    def result_derivs(self, slot):
        """(value, d/dmean_x, d/dmean_y) of a mixture slot with derivatives."""
        w, hh = self._slots[slot]
        if w * hh == 0:
            z = np.zeros((max(hh, 0), max(w, 0)), np.float32)
            return z, z.copy(), z.copy()
        o = int(self._off[slot])
        sl = slice(o, o + w * hh)
        return (self._host[sl].reshape(hh, w).copy(), self._dx[sl].reshape(hh, w).copy(),
                self._dy[sl].reshape(hh, w).copy())


def resolve(fn, *args, **kwargs):
    """Plan with fn(plan, *args, **kwargs), run, and return the Pending's
    value: the single-call form of every planned render."""
    plan = RenderPlan()
    p = fn(plan, *args, **kwargs)
    plan.run()
    return p.get()
