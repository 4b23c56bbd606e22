"""The oracle's EM fit of a mixture of Gaussians to a PSF stamp
(em_fit_2d_reg, emfit.i:274; GaussianMixturePSF.fromStamp, psf.py:537-597)
pinned by closed forms: one component reaches the image's centroid and
second moments (plus the alpha prior), two well-separated blobs are each
recovered by one component, and pixels beyond Mahalanobis distance 10 of
every component are ignored."""
import numpy as np

import oracle as O


def _gauss_img(NX, NY, x0, y0, mx, my, cov, amp=1.):
    yy, xx = np.mgrid[0:NY, 0:NX]
    d = np.stack([x0 + xx - mx, y0 + yy - my], -1)
    ic = np.linalg.inv(cov)
    q = np.einsum('...i,ij,...j->...', d, ic, d)
    return amp * np.exp(-0.5 * q)


def _moments(img, x0, y0):
    NY, NX = img.shape
    yy, xx = np.mgrid[0:NY, 0:NX]
    X, Y = x0 + xx, y0 + yy
    w = img / img.sum()
    mx, my = np.sum(w * X), np.sum(w * Y)
    c = np.array([[np.sum(w * (X - mx) ** 2), np.sum(w * (X - mx) * (Y - my))],
                  [np.sum(w * (X - mx) * (Y - my)), np.sum(w * (Y - my) ** 2)]])
    return mx, my, c


def test_one_component_reaches_moments():
    img = _gauss_img(21, 17, -10, -8, 0.7, -0.4, np.array([[3., 0.8], [0.8, 2.]]))
    img /= img.sum()
    mx, my, c = _moments(img, -10, -8)
    # one step from a start whose support covers the stamp: Z = 1 everywhere
    amp, mean, var = O.em_fit_2d_reg(img, -10, -8, [1.], [[0., 0.]], [np.eye(2) * 40.], 0., 1)
    assert abs(amp[0] - 1.) < 1e-12
    np.testing.assert_allclose(mean[0], [mx, my], rtol=0, atol=1e-12)
    np.testing.assert_allclose(var[0], c, rtol=1e-12)
    # the alpha prior: (alpha I + wsum c) / (wsum + alpha)
    alpha = 0.3
    _, _, var_a = O.em_fit_2d_reg(img, -10, -8, [1.], [[0., 0.]], [np.eye(2) * 40.], alpha, 1)
    np.testing.assert_allclose(var_a[0], (alpha * np.eye(2) + c) / (1. + alpha), rtol=1e-12)


def test_two_separated_blobs():
    c1 = np.array([[2., 0.3], [0.3, 1.5]])
    c2 = np.array([[1.2, -0.2], [-0.2, 2.5]])
    a = _gauss_img(60, 30, -30, -15, -14., 0., c1, 0.6)
    b = _gauss_img(60, 30, -30, -15, 15., 1., c2, 0.4)
    img = a + b
    s = img.sum()
    amp, mean, var = O.em_fit_2d_reg(img / s, -30, -15, [0.5, 0.5], [[-10., 0.], [10., 0.]],
                                     [np.eye(2) * 4., np.eye(2) * 4.], 0., 200)
    for k, blob in enumerate((a, b)):
        mx, my, c = _moments(blob, -30, -15)
        assert abs(amp[k] - blob.sum() / s) < 1e-6
        np.testing.assert_allclose(mean[k], [mx, my], atol=1e-5)
        np.testing.assert_allclose(var[k], c, atol=1e-4)


def test_far_pixels_are_ignored():
    img = np.zeros((41, 41))
    img[20, 20] = 1.
    img[19:22, 19:22] += 0.1
    img_far = img.copy()
    img_far[0, 0] = 5.            # 20 px away, sigma 1: distance^2 800 >= 100
    r1 = O.em_fit_2d_reg(img, -20, -20, [1.], [[0., 0.]], [np.eye(2)], 0., 1)
    r2 = O.em_fit_2d_reg(img_far, -20, -20, [1.], [[0., 0.]], [np.eye(2)], 0., 1)
    for u, v in zip(r1, r2):
        np.testing.assert_allclose(u, v, rtol=1e-14)
