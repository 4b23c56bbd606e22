"""Pin the oracle's ring-growth MoG evaluator (c_gauss_2d_approx3,
tractor/approx3.c) with invariants the algorithm fixes (CPU only).  The
reference's own self-test only prints approx3's difference from the exact grid
(tractor/mixture_profiles.py:690-697); these checks are what the mathematics
and the growth rule determine."""
import math

import numpy as np

import oracle as O


def _mix(K=2, seed=0):
    rng = np.random.RandomState(seed)
    amp = rng.uniform(0.5, 2., K)
    mean = rng.uniform(-0.5, 0.5, (K, 2))
    var = np.zeros((K, 2, 2))
    for k in range(K):
        a, d = rng.uniform(2., 6., 2)
        b = rng.uniform(-0.3, 0.3) * math.sqrt(a * d)
        var[k] = [[a, b], [b, d]]
    return amp, mean, var


def _exact(x0, x1, y0, y1, fx, fy, amp, mean, var):
    return O.mog_grid(x0, x1, y0, y1, fx, fy, amp, mean, var)


def test_minval_zero_evaluates_whole_box_exactly():
    # minval 0: every ring pixel marks its neighbours, so the whole box is
    # evaluated and equals the exact grid (floor -100 on the folded exponent
    # is exp(-100) relative: invisible).
    amp, mean, var = _mix(3, 1)
    x0, x1, y0, y1 = -12, 15, -10, 9
    fx, fy = 0.3, -0.7
    res, xd, yd, bb = O.mog_grid_approx3(x0, x1, y0, y1, fx, fy, 0., amp, mean, var, 0, 0)
    ex = _exact(x0, x1, y0, y1, fx, fy, amp, mean, var)
    assert bb == (0, x1 - x0, 0, y1 - y0)
    # the grid drops a component at Mahalanobis d >= 100 (exp(-50) of its
    # peak), approx3 only below d = 200: the difference is below 1e-20.
    np.testing.assert_allclose(res, ex, rtol=1e-13, atol=1e-20 * ex.max())


def test_derivatives_are_d_dmean():
    # xderiv / yderiv are derivatives with respect to the component means:
    # moving every mean by +h equals moving the evaluation offset (fx, fy).
    amp, mean, var = _mix(2, 2)
    box = (-8, 9, -7, 8)
    res, xd, yd, _ = O.mog_grid_approx3(*box, 0.1, 0.2, 0., amp, mean, var, 0, 0)
    h = 1e-5
    ep = _exact(*box, 0.1 + h, 0.2, amp, mean, var)
    em = _exact(*box, 0.1 - h, 0.2, amp, mean, var)
    np.testing.assert_allclose(xd, (ep - em) / (2 * h), rtol=1e-5, atol=1e-10)
    ep = _exact(*box, 0.1, 0.2 + h, amp, mean, var)
    em = _exact(*box, 0.1, 0.2 - h, amp, mean, var)
    np.testing.assert_allclose(yd, (ep - em) / (2 * h), rtol=1e-5, atol=1e-10)


def test_truncation_region_single_gaussian():
    # One isotropic Gaussian: every pixel at or above minval is reached (the
    # superlevel set is a disk around the start pixel), values there equal the
    # exact ones, nothing is evaluated more than one pixel beyond the set
    # (outside the first minradius rings), and the bbox is tight.
    amp = np.array([1.])
    mean = np.zeros((1, 2))
    var = np.array([np.eye(2) * 9.])
    x0, x1, y0, y1 = -30, 31, -30, 31
    minval = 1e-4
    res, _, _, bb = O.mog_grid_approx3(x0, x1, y0, y1, 0.2, -0.1, minval, amp, mean, var, 0, 0)
    ex = _exact(x0, x1, y0, y1, 0.2, -0.1, amp, mean, var)
    above = ex >= minval
    assert np.all(res[above] > 0)
    ev = res > 0
    np.testing.assert_allclose(res[ev], ex[ev], rtol=1e-13)
    # dilate 'above' by one pixel (8-neighbourhood)
    dil = above.copy()
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            dil |= np.roll(np.roll(above, dy, 0), dx, 1)
    assert not np.any(ev & ~dil)
    # bbox covers the evaluated pixels; the outermost evaluated ring can be
    # all zeros (below every component floor), so it may exceed the non-zero
    # bounds by one pixel per side.
    ys, xs = np.nonzero(ev)
    nz = (xs.min(), xs.max() + 1, ys.min(), ys.max() + 1)
    assert nz[0] - 1 <= bb[0] <= nz[0] and nz[2] - 1 <= bb[2] <= nz[2]
    assert nz[1] <= bb[1] <= nz[1] + 1 and nz[3] <= bb[3] <= nz[3] + 1
    # a tail pixel 15 px out (value ~7e-8 < minval) is not evaluated
    assert res[30, 45] == 0 and 0 < ex[30, 45] < minval


def test_component_floor():
    # maxD = log(minval 2 pi sqrt(det) / sum amp): a pixel where one component's
    # exponent is below its floor gets only the other components.
    amp = np.array([1., 1.])
    mean = np.array([[0., 0.], [6., 0.]])
    var = np.array([np.eye(2), np.eye(2)])
    minval = 1e-3
    res, _, _, _ = O.mog_grid_approx3(-3, 10, -3, 4, 0., 0., minval, amp, mean, var, 0, 0)
    md = math.log(minval * 2 * math.pi * 1. / 2.)
    # pixel (x=3, y=0): both exponents -4.5 >= md (~ -5.76): both counted
    v = res[3, 3 + 3]
    assert abs(v - 2 * math.exp(-4.5) / (2 * math.pi)) < 1e-15
    # pixel (x=0, y=0): component 2 exponent -18 < md: skipped
    assert abs(res[3, 3] - 1. / (2 * math.pi)) < 1e-15
    assert md > -18.


def test_start_outside_box_uses_brightest_edge_pixel():
    # Source left of the box: growth starts at the brightest pixel of the left
    # edge, and the evaluated region touches that edge.
    amp, mean, var = _mix(1, 3)
    res, _, _, bb = O.mog_grid_approx3(5, 25, -6, 7, 0., 0., 1e-3, amp, mean, var, 0, 0)
    ex = _exact(5, 25, -6, 7, 0., 0., amp, mean, var)
    iy = int(np.argmax(ex[:, 0]))
    assert res[iy, 0] == ex[iy, 0] and bb[0] == 0
    # everything evaluated matches the exact value to the component floor
    ev = res > 0
    assert np.all(np.abs(res[ev] - ex[ev]) <= 1e-3)


def test_minradius_rings_always_grow():
    # With minval above the peak nothing passes the threshold (and every
    # value is floored to 0), but the first minradius rings still grow: the
    # evaluated box is the (2m+3)^2 square around the start pixel.
    amp = np.array([1.])
    mean = np.zeros((1, 2))
    var = np.array([np.eye(2) * 4.])
    for m in (0, 1, 3):
        res, _, _, bb = O.mog_grid_approx3(-20, 21, -20, 21, 0., 0., 10., amp, mean, var, 0, 0,
                                           minradius=m)
        assert bb == (20 - (m + 1), 20 + m + 2, 20 - (m + 1), 20 + m + 2), (m, bb)
        assert np.count_nonzero(res) == 0
