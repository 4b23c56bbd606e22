"""Chained restarts in the batched fitter (trh_fitter_set_chain): the
second source of a pair waits, stopped, until its first has stopped, and
then runs from its own start only if the first's chi^2 is above the pair's
reference (or not a number).  A restarted source's fit equals the fit of the
same source from the same start in a batch of its own, bit for bit (every
source's fit is independent of the others in its batch), and so does every
first source's; the chain launch is part of the captured graph."""
import numpy as np
import pytest
import torch

from tractor_amd import _lib, batch as B, workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def scene():
    assert torch.cuda.is_available()
    n = 48
    kind, P, truth = W.draw_sources(n, rng=np.random.default_rng(W.seed(4, 60)),
                                    P_choices=np.array([24, 40, 56, 72]))
    psf = W.moffat_stamp(63, 4.0, 3.5)
    f, data, inverr, pix_off = B.make_problem(kind, P, truth, psf, W.seed(4, 61))
    f.close()
    s0 = W.perturb(kind, truth, np.random.default_rng(W.seed(4, 62)))
    s1 = W.perturb(kind, truth, np.random.default_rng(W.seed(4, 63)))
    return dict(n=n, kind=kind, P=P, psf=psf, data=data, inverr=inverr, pix_off=pix_off,
                s0=s0, s1=s1)


def _alone(sc, start, graph=True):
    f = B.BatchFitter(sc['kind'], sc['P'], sc['pix_off'], sc['data'], sc['inverr'], sc['psf'])
    f.set_params(start)
    f.fit(max_iter=100, graph=graph)
    st, p = f.status(), f.get_params()
    f.close()
    return st, p


def _chained(sc, ref, graph=True):
    n = sc['n']
    f = B.BatchFitter(np.tile(sc['kind'], 2), np.tile(sc['P'], 2), np.tile(sc['pix_off'], 2),
                      sc['data'], sc['inverr'], sc['psf'])
    f.set_chain(np.arange(n), n + np.arange(n), ref)
    f.set_params(np.concatenate([sc['s1'], sc['s0']]))
    f.fit(max_iter=100, graph=graph)
    st, p, cs = f.status(), f.get_params(), f.chain_state()
    return f, st, p, cs


def test_restarts_follow_the_rule_and_equal_separate_fits(scene):
    sc = scene
    n = sc['n']
    print('chained restarts vs separate fits', flush=True)
    a0, p0 = _alone(sc, sc['s0'])
    a1, p1 = _alone(sc, sc['s1'])
    # pairs 0, 3, 6, ...: reference 0 (always above: restart); 1, 4, ...:
    # +inf (never); 2, 5, ...: NaN (not a number: restart)
    ref = np.where(np.arange(n) % 3 == 0, 0.0, np.where(np.arange(n) % 3 == 1, np.inf, np.nan))
    f, st, p, cs = _chained(sc, ref)
    go = np.arange(n) % 3 != 1
    np.testing.assert_array_equal(cs, np.where(go, 1, 2))
    # first sources: as their own fit from s1
    np.testing.assert_array_equal(p[:n], p1)
    np.testing.assert_array_equal(st['chisq'][:n], a1['chisq'])
    np.testing.assert_array_equal(st['iterations'][:n], a1['iterations'])
    # restarted seconds: as their own fit from s0; the others never ran
    np.testing.assert_array_equal(p[n:][go], p0[go])
    np.testing.assert_array_equal(st['chisq'][n:][go], a0['chisq'][go])
    np.testing.assert_array_equal(st['iterations'][n:][go], a0['iterations'][go])
    assert np.all(st['iterations'][n:][~go] == 0)
    np.testing.assert_array_equal(p[n:][~go], sc['s0'][~go][:, :p.shape[1]])
    assert st['n_active'] == 0
    # set_params makes every pair pending again: the same fit once more
    f.set_params(np.concatenate([sc['s1'], sc['s0']]))
    np.testing.assert_array_equal(f.chain_state(), np.zeros(n, np.int32))
    f.fit(max_iter=100)
    np.testing.assert_array_equal(f.get_params(), p)
    np.testing.assert_array_equal(f.chain_state(), cs)
    f.close()


def test_eager_equals_graph_and_removal(scene):
    sc = scene
    n = sc['n']
    print('chained restarts: eager vs graph', flush=True)
    ref = np.where(np.arange(n) % 2 == 0, 0.0, np.inf)
    fg, sg, pg, cg = _chained(sc, ref, graph=True)
    fg.close()
    fe, se, pe, ce = _chained(sc, ref, graph=False)
    np.testing.assert_array_equal(pe, pg)
    np.testing.assert_array_equal(ce, cg)
    np.testing.assert_array_equal(se['iterations'], sg['iterations'])
    # n = 0 removes the pairs: every source runs from the next set_params on
    fe.set_chain([], [], [])
    fe.set_params(np.concatenate([sc['s1'], sc['s0']]))
    fe.fit(max_iter=100)
    st = fe.status()
    assert np.all(st['iterations'] > 0)
    fe.close()


def test_bad_pairs_are_rejected(scene):
    sc = scene
    n = sc['n']
    f = B.BatchFitter(sc['kind'], sc['P'], sc['pix_off'], sc['data'], sc['inverr'], sc['psf'])
    ref = np.zeros(2)
    for first, second in (([0, 1], [2, 2]),      # a source in two pairs
                          ([0, 1], [1, 3]),      # a source in both roles
                          ([0, 1], [0, 3]),      # first == second
                          ([0, 1], [2, n]),      # out of range
                          ([-1, 1], [2, 3])):
        with pytest.raises(RuntimeError):
            f.set_chain(first, second, ref)
    with pytest.raises(RuntimeError):
        f.set_chain(np.arange(n // 2 + 1), np.arange(n // 2 + 1) + n // 2, np.zeros(n // 2 + 1))
    f.close()
