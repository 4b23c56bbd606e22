"""Pin the CNN oracle (CPU only): its convolution against scipy's 2-D
correlation, its analytic gradients against central finite differences of its
own FP64 loss, and the bf16 rounding rule against exact cases."""
import numpy as np
import pytest

from oracle import cnn_oracle as N


def _params(rng, scale=0.3):
    P = {}
    for l, cin in enumerate([1, 32, 32, 32]):
        P['cw%d' % l] = rng.normal(0, scale * np.sqrt(2. / (9 * cin)), (32, 9, cin))
        P['cb%d' % l] = rng.normal(0, 0.05, 32)
    P['hW1'] = rng.normal(0, np.sqrt(2. / 32), (64, 32))
    P['hb1'] = rng.normal(0, 0.05, 64)
    P['hW2'] = rng.normal(0, np.sqrt(1. / 64), (9, 64))
    P['hb2'] = rng.normal(0, 0.05, 9)
    return P


def test_bf16_round_exact_cases():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 0.0, 3.0e38], np.float32)
    r = N.bf16_round(x)
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-8 ties up to 1 + 2^-6
    assert list(r[:5]) == [1.0, 1.0, 1.0 + 2 ** -6, -2.5, 0.0]
    assert np.isfinite(r[5])


def test_conv_matches_scipy_correlate():
    from scipy.signal import correlate2d
    rng = np.random.RandomState(0)
    x = rng.normal(size=(2, 10, 12, 3))
    w = rng.normal(size=(4, 9, 3))
    b = rng.normal(size=4)
    z = N.conv_fwd(x, w, b)
    for bi in range(2):
        for co in range(4):
            ref = np.full((10, 12), b[co])
            for ci in range(3):
                k = w[co, :, ci].reshape(3, 3)
                ref += correlate2d(x[bi, :, :, ci], k, mode='same')
            np.testing.assert_allclose(z[bi, :, :, co], ref[::2, ::2], rtol=1e-12, atol=1e-12)


def test_gradients_match_finite_differences():
    rng = np.random.RandomState(1)
    B = 3
    x = rng.uniform(0, 1, (B, 32, 32, 1))
    P = _params(rng, scale=1.0)
    T = rng.normal(size=(B, 6))
    lab = np.array([0, 1, 2])
    R = N.forward_backward(x, P, T, lab, bf16=False)
    h = 1e-6
    for key, n in [('cw0', 5), ('cb0', 3), ('cw1', 5), ('cw2', 5), ('cw3', 5), ('cb3', 3),
                   ('hW1', 4), ('hb1', 3), ('hW2', 4), ('hb2', 3)]:
        g = R['g_' + key]
        flat = P[key].reshape(-1)
        for idx in rng.choice(flat.size, n, replace=False):
            old = flat[idx]
            flat[idx] = old + h
            lp = N.forward_backward(x, P, T, lab, bf16=False)['loss']
            flat[idx] = old - h
            lm = N.forward_backward(x, P, T, lab, bf16=False)['loss']
            flat[idx] = old
            fd = (lp - lm) / (2 * h)
            an = g.reshape(-1)[idx]
            assert abs(fd - an) <= 1e-6 + 1e-5 * abs(fd), (key, idx, fd, an)


def test_point_sources_ignore_shape_targets():
    rng = np.random.RandomState(2)
    x = rng.uniform(0, 1, (2, 32, 32, 1))
    P = _params(rng)
    T = rng.normal(size=(2, 6))
    R0 = N.forward_backward(x, P, T, np.array([0, 0]), bf16=False)
    T2 = T.copy()
    T2[:, 3:] += 5.
    R1 = N.forward_backward(x, P, T2, np.array([0, 0]), bf16=False)
    assert R0['loss'] == R1['loss']
