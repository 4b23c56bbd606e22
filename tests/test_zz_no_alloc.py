"""trh_lanczos_shift_batch_noalloc (the allocation-free entry point) against
trh_lanczos_shift_batch: the same kernels, so the outputs must be identical
bit for bit, on two calls in a row (the second smaller, so stale contents of
the shared scratch array would show), plus the host-side size check."""
import sys

import numpy as np
import pytest
import torch

from tractor_amd import _lib
from tractor_amd._lib import TRH_NOALLOC_SCRATCH_FLOATS, check, ptr

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _shift_both(lib, h, imgs, dx, dy):
    n, H, W = imgs.shape
    din = torch.from_numpy(imgs.ravel().copy()).cuda()
    ref = torch.full_like(din, float('nan'))
    got = torch.full_like(din, float('nan'))
    ddx = torch.from_numpy(dx).cuda()
    ddy = torch.from_numpy(dy).cuda()
    check(lib.trh_lanczos_shift_batch(h.value, ptr(din), ptr(ref), n, H, W, ptr(ddx), ptr(ddy),
                                      None), 'trh_lanczos_shift_batch')
    check(lib.trh_lanczos_shift_batch_noalloc(h.value, ptr(din), ptr(got), n, H, W, ptr(ddx),
                                              ptr(ddy), None), 'trh_lanczos_shift_batch_noalloc')
    torch.cuda.synchronize()
    return ref.cpu().numpy(), got.cpu().numpy()


def test_noalloc_matches_allocating_entry_point():
    assert torch.cuda.is_available()
    lib = _lib.load()
    h = _lib.handle(0)
    rng = np.random.default_rng(20261204)
    # first call: 3 images of 200 x 300; second: 2 images of 64 x 48
    for n, H, W in [(3, 200, 300), (2, 64, 48)]:
        say('noalloc call', n, H, W)
        imgs = rng.normal(size=(n, H, W)).astype(np.float32)
        dx = rng.uniform(-0.5, 0.5, n)
        dy = rng.uniform(-0.5, 0.5, n)
        ref, got = _shift_both(lib, h, imgs, dx, dy)
        assert np.all(np.isfinite(ref))
        assert np.array_equal(ref, got)


def test_noalloc_rejects_oversize_without_running():
    lib = _lib.load()
    h = _lib.handle(0)
    x = torch.zeros(16, dtype=torch.float32, device='cuda')
    d = torch.zeros(1, dtype=torch.float64, device='cuda')
    side = 4096                                    # 16.7 M floats > 8.4 M
    assert side * side > TRH_NOALLOC_SCRATCH_FLOATS
    rc = lib.trh_lanczos_shift_batch_noalloc(h.value, ptr(x), ptr(x), 1, side, side, ptr(d),
                                             ptr(d), None)
    assert rc == -1
    assert b'scratch' in lib.trh_last_error()
