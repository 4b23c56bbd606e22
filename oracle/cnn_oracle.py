"""NumPy restatement of the surrogate CNN (forward, loss and backward).

TEST INFRASTRUCTURE ONLY: the checker of tractor_amd's device CNN kernels;
nothing in tractor_amd imports it.  The network has no counterpart in the
reference (it is the learned initialiser of the batched fit), so this module
is pinned by the mathematics instead: finite-difference gradients and an
independent 2-D correlation (tests/test_oracle_cnn.py).

Layout (matches the device): activations NHWC; conv weights cw<l> of shape
(32, 9, cin) = [co][tap = 3*ky + kx][ci], biases cb<l> (32,); head hW1
(64, 32), hb1 (64,), hW2 (9, 64), hb2 (9,).  Convolutions are 3x3, stride 2, pad 1, ReLU.
With ``bf16=True`` every tensor the device stores in bfloat16 (weights,
activations, pre-activation gradients) is rounded the same way.
"""
import numpy as np


def bf16_round(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7fff + ((u >> 16) & 1)) & 0xffff0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def _r(a, on):
    return bf16_round(a).astype(np.float64) if on else np.asarray(a, np.float64)


def conv_fwd(x, w, b):
    """x (B, H, W, cin), w (32, 9, cin), b (32,) -> pre-activation (B, H/2, W/2, 32)."""
    B, H, W, cin = x.shape
    OH, OW = H // 2, W // 2
    xp = np.zeros((B, H + 2, W + 2, cin))
    xp[:, 1:H + 1, 1:W + 1] = x
    out = np.zeros((B, OH, OW, w.shape[0])) + b
    for ky in range(3):
        for kx in range(3):
            patch = xp[:, ky:ky + 2 * OH:2, kx:kx + 2 * OW:2, :]
            out += patch @ w[:, 3 * ky + kx, :].T
    return out


def conv_bwd(x, w, dz):
    """Gradients of sum(dz * conv_fwd(x, w, b)): (dx, dw, db)."""
    B, H, W, cin = x.shape
    OH, OW = H // 2, W // 2
    xp = np.zeros((B, H + 2, W + 2, cin))
    xp[:, 1:H + 1, 1:W + 1] = x
    dxp = np.zeros_like(xp)
    dw = np.zeros_like(w)
    for ky in range(3):
        for kx in range(3):
            t = 3 * ky + kx
            patch = xp[:, ky:ky + 2 * OH:2, kx:kx + 2 * OW:2, :]
            dw[:, t, :] = np.einsum('bhwo,bhwi->oi', dz, patch)
            dxp[:, ky:ky + 2 * OH:2, kx:kx + 2 * OW:2, :] += dz @ w[:, t, :]
    return dxp[:, 1:H + 1, 1:W + 1], dw, dz.sum(axis=(0, 1, 2))


def forward_backward(x, P, targets=None, labels=None, rw=None, tw=1.0, bf16=True):
    """x (B, 32, 32, 1); P dict with cw0..cw3, cb0..cb3, hW1, hb1, hW2, hb2.
    Returns dict with activations y1..y4, out (B, 9), and with targets: loss,
    dz4 (pre-activation gradient of layer 4), dz3..dz1 and all parameter
    gradients (g_cw0.., g_cb0.., g_hW1, g_hb1, g_hW2, g_hb2)."""
    R = {}
    y = _r(x, bf16)
    ys = [y]
    zs = []
    for l in range(4):
        w = _r(P['cw%d' % l], bf16)
        z = conv_fwd(y, w, np.asarray(P['cb%d' % l], np.float64))
        y = _r(np.maximum(z, 0.).astype(np.float32), bf16) if bf16 else np.maximum(z, 0.)
        zs.append(z)
        ys.append(y)
        R['y%d' % (l + 1)] = y
    y4 = ys[4]
    Bn = x.shape[0]
    f = y4.mean(axis=(1, 2))
    W1, b1, W2, b2 = (np.asarray(P[k], np.float64) for k in ('hW1', 'hb1', 'hW2', 'hb2'))
    z1 = f @ W1.T + b1
    a1 = np.maximum(z1, 0.)
    o = a1 @ W2.T + b2
    R['out'] = o
    if targets is None:
        return R
    rw = np.ones(6) if rw is None else np.asarray(rw, np.float64)
    lab = np.asarray(labels)
    wgt = np.tile(rw, (Bn, 1))
    wgt[lab == 0, 3:] = 0.
    e = o[:, :6] - targets
    loss = 0.5 * np.sum(wgt * e * e)
    lg = o[:, 6:]
    lg = lg - lg.max(axis=1, keepdims=True)
    p = np.exp(lg)
    p /= p.sum(axis=1, keepdims=True)
    loss += -tw * np.sum(np.log(np.maximum(p[np.arange(Bn), lab], 1e-30)))
    R['loss'] = loss / Bn
    d = np.zeros_like(o)
    d[:, :6] = wgt * e / Bn
    oh = np.zeros_like(p)
    oh[np.arange(Bn), lab] = 1.
    d[:, 6:] = tw * (p - oh) / Bn
    R['g_hW2'] = d.T @ a1
    R['g_hb2'] = d.sum(0)
    dz1 = (d @ W2) * (z1 > 0)
    R['g_hW1'] = dz1.T @ f
    R['g_hb1'] = dz1.sum(0)
    df = dz1 @ W1
    dz = np.repeat(np.repeat(df[:, None, None, :] * 0.25, 2, 1), 2, 2) * (y4 > 0)
    if bf16:
        dz = _r(dz.astype(np.float32), True)
    R['dz4'] = dz
    for l in range(3, -1, -1):
        w = _r(P['cw%d' % l], bf16)
        dx, dw, db = conv_bwd(ys[l], w, dz)
        R['g_cw%d' % l] = dw
        R['g_cb%d' % l] = db
        if l > 0:
            dz = dx * (ys[l] > 0)
            if bf16:
                dz = _r(dz.astype(np.float32), True)
            R['dz%d' % l] = dz
    return R
