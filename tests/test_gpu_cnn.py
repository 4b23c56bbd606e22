"""Surrogate CNN on the device against the NumPy oracle (oracle/cnn_oracle.py):
forward activations, loss, pre-activation gradients and every parameter
gradient of the BF16-MFMA kernels; then a short training run on cutouts from
the library's own renderer, and what the trained network gains over the plain
start.

Tolerances: the device stores weights, activations and pre-activation
gradients in bfloat16 (8-bit mantissa) exactly where the oracle rounds them,
so the only differences are FP32 versus FP64 accumulation order and the rare
bfloat16 rounding flips they cause; those stay below 1e-2 in relative norm."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module', autouse=True)
def _need_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')


def _rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


def _setup(B=16, seed=3):
    import torch
    from tractor_amd import surrogate as S
    rng = np.random.default_rng(seed)
    p = S.init_params(seed)
    p[S.OFF_B[0]:S.OFF_B[0] + 32] = rng.normal(0, 0.05, 32)
    net = S.SurrogateCNN(params=p, max_batch=B)
    x = rng.uniform(-0.2, 1.0, (B, 32, 32, 1)).astype(np.float32)
    X = torch.from_numpy(x).to('cuda').to(torch.bfloat16)
    T = rng.normal(size=(B, 6)).astype(np.float32)
    L = rng.integers(0, 3, B).astype(np.int32)
    return net, X, T, L, p


def test_forward_matches_oracle():
    import torch
    from oracle import cnn_oracle as N
    from tractor_amd import surrogate as S
    net, X, T, L, p = _setup()
    B = X.shape[0]
    print('cnn forward B=%d' % B, flush=True)
    net.pack()
    net.forward(X, B)
    net.head(B)
    torch.cuda.synchronize()
    xr = X.float().cpu().numpy()
    R = N.forward_backward(xr, S.unpack_params(p), bf16=True)
    sizes = (16, 8, 4, 2)
    for l in range(4):
        y = net.y[l][:B * sizes[l] ** 2 * 32].float().cpu().numpy().reshape(B, sizes[l], sizes[l], 32)
        assert _rel(y, R['y%d' % (l + 1)]) < 1e-2, l
    out = net.out[:B * 9].cpu().numpy().reshape(B, 9)
    assert _rel(out, R['out']) < 1e-2


def test_backward_matches_oracle():
    import torch
    from oracle import cnn_oracle as N
    from tractor_amd import surrogate as S
    net, X, T, L, p = _setup(B=24, seed=5)
    B = X.shape[0]
    print('cnn backward B=%d' % B, flush=True)
    dT = torch.from_numpy(T).to('cuda')
    dL = torch.from_numpy(L).to('cuda')
    rw = torch.ones(6, dtype=torch.float32, device='cuda')
    net.g.zero_()
    net.loss.zero_()
    net.pack()
    net.forward(X, B)
    net.head(B, dT, dL, rw, 1.0)
    net.backward(X, B)
    torch.cuda.synchronize()
    R = N.forward_backward(X.float().cpu().numpy(), S.unpack_params(p), T, L, bf16=True)
    assert abs(float(net.loss.item()) - R['loss']) < 1e-3 * abs(R['loss'])
    sizes = (16, 8, 4, 2)
    for l in (3, 2, 1):
        dz = net.dz[l][:B * sizes[l] ** 2 * 32].float().cpu().numpy()
        assert _rel(dz, R['dz%d' % (l + 1)]) < 2e-2, l
    G = S.unpack_params(net.g.cpu().numpy())
    for k in G:
        assert _rel(G[k], R['g_' + k]) < 2e-2, (k, _rel(G[k], R['g_' + k]))


def test_training_learns_and_helps_the_fit():
    """Train on 262144 rendered cutouts for 4 epochs; the loss falls, types
    are recovered well above the 40% majority class, positions land far closer
    than the standard start, and on a fresh 2048-source problem the
    surrogate-started fit (with the self-check fallback) ends no worse than
    1e-4 in chi^2 than the plain fit for most sources (about 90% on the
    16384-source benchmark: the rest are mostly exp/dev type swaps), and with
    second-type hypotheses below 0.7 confidence for more of them; with the
    fallback against the plain fit's chi^2 (re-fit when 1e-4 worse), none."""
    import torch
    from tractor_amd import surrogate as S, workloads as W, batch as Bt
    print('cnn training set', flush=True)
    X, T, L = S.make_training_set(262144, chunk=16384, seed=11)
    net = S.SurrogateCNN(max_batch=1024, seed=1)
    hist = net.fit(X, T, L, epochs=4, batch=1024, lr=3e-3)
    print('loss history', hist[0], hist[-1], flush=True)
    assert hist[0][0] == 1 and hist[-1][1] < 0.4 * hist[0][1]
    del X, T, L
    n = 2048
    rng = np.random.default_rng(99)
    kind, P, truth = W.draw_sources(n, rng=rng)
    psf = W.moffat_stamp(63, 4.0, 3.5)
    fitter, data, inverr, pix_off = Bt.make_problem(kind, P, truth, psf, 1234)
    Xh, maxv = S.cutouts(data, pix_off, P)
    pk, start = net.decode(net.predict(Xh, n), maxv.cpu().numpy(), P)
    acc = np.mean(pk == kind)
    plain_start = W.perturb(kind, truth, np.random.default_rng(5))
    perr = np.median(np.abs(start[:, 0] - truth[:, 0]))
    serr = np.median(np.abs(plain_start[:, 0] - truth[:, 0]))
    print('type accuracy %.3f; median |x err| proposal %.3f, standard start %.3f' % (acc, perr, serr),
          flush=True)
    assert acc > 0.7
    assert perr < 0.3 * serr
    fitter.set_params(plain_start)
    fitter.fit(max_iter=100)
    plain = fitter.status()
    fitter.close()
    fracs = []
    for tc in (0.0, 0.7):
        res = S.fit_with_fallback(net, data, inverr, pix_off, P, psf, plain_start, plain_kind=kind,
                                  type_conf=tc)
        rel = (res['chisq'] - plain['chisq']) / plain['chisq']
        frac = np.mean(rel <= 1e-4)
        fracs.append(frac)
        print('surrogate type_conf %.1f: frac not worse than plain chi2 by 1e-4: %.4f; %s; '
              'iterations %.2f vs %.2f' % (tc, frac, res['timings'], res['iterations'].mean(),
                                           plain['iterations'].mean()), flush=True)
    assert fracs[0] > 0.85
    assert fracs[1] > 0.93 and fracs[1] >= fracs[0]
    # the specified rule: re-fit from the standard start every source whose
    # final chi^2 exceeds the plain fit's by more than 1e-4 relative; the
    # re-fit reproduces the plain fit, so no source ends worse than that
    res = S.fit_with_fallback(net, data, inverr, pix_off, P, psf, plain_start, plain_kind=kind,
                              plain_chisq=plain['chisq'])
    rel = (res['chisq'] - plain['chisq']) / plain['chisq']
    print('plain-chi2 fallback: %d sources re-fit; worst rel chi2 vs plain %.3e; %s' % (
        int(res['fallback'].sum()), rel.max(), res['timings']), flush=True)
    assert np.all(rel <= 1e-4)
    assert res['fallback'].sum() == res['timings']['fallback_sources'] < 0.2 * n
    # the re-fits as a separate batch after the fit (the default) and as
    # chained restarts in the same batch give the same results: every
    # source's fit is independent of the others in its batch
    assert not res['timings']['fallback_overlapped']
    ovl = S.fit_with_fallback(net, data, inverr, pix_off, P, psf, plain_start, plain_kind=kind,
                              plain_chisq=plain['chisq'], overlap=True)
    assert ovl['timings']['fallback_overlapped']
    print('sequential fallback %.1f ms, overlapped %.1f ms' % (
        1e3 * res['timings']['total_s'], 1e3 * ovl['timings']['total_s']), flush=True)
    np.testing.assert_array_equal(ovl['fallback'], res['fallback'])
    np.testing.assert_array_equal(ovl['chisq'], res['chisq'])
    np.testing.assert_array_equal(ovl['params'], res['params'])
    np.testing.assert_array_equal(ovl['kind'], res['kind'])
    np.testing.assert_array_equal(ovl['iterations'], res['iterations'])
