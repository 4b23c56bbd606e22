"""SersicMixture (tractor/sersic.py:25-341) and the Sersic parameters, on
the host.  The table's nodes are reproduced exactly by the interpolation;
at n = 1 and n = 4 the mixtures agree with the separately tabulated exp and
dev profiles (mixture_profiles.py:17-38) to their fitting precision."""
import numpy as np
import pytest

from tractor_amd import mixture_profiles as mp
from tractor_amd.sersic import SersicGalaxy, SersicIndex, SersicMixture
from tractor_amd.sersic_table import CORES, RANGES


def _node(n):
    for lo, hi, nodes in RANGES:
        for s, a, v in nodes:
            if s == n and lo <= n < hi:
                return np.array(a), np.array(v)
    raise KeyError(n)


@pytest.mark.parametrize('n', [0.3, 0.45, 0.8, 2.0, 5.0])
def test_profile_reproduces_table_nodes(n):
    a, v = _node(n)
    a = a / a.sum()
    if n > 1.:
        core = dict(CORES).get(n)
        if core is None:
            core = float(SersicMixture.getProfile(n).amp[0])   # spline between core nodes
        a = np.append(a * (1. - core), core)
        v = np.append(v, 0.)
    o = np.argsort(v)
    m = SersicMixture.getProfile(n)
    np.testing.assert_allclose(m.amp, a[o], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(m.var[:, 0, 0], v[o], rtol=1e-12)
    np.testing.assert_allclose(m.var[:, 1, 1], v[o], rtol=1e-12)
    assert np.all(m.var[:, 0, 1] == 0.) and np.all(m.mean == 0.)


@pytest.mark.parametrize('n,ref', [(1.0, mp.get_exp_mixture), (4.0, mp.get_dev_mixture)])
def test_exp_and_dev_special_cases(n, ref):
    m = SersicMixture.getProfile(n)
    r = ref()
    o = np.argsort(r.var[:, 0, 0])
    assert m.K == r.K
    np.testing.assert_allclose(m.amp, r.amp[o], rtol=1e-3, atol=2e-6)
    np.testing.assert_allclose(m.var[:, 0, 0], r.var[o, 0, 0], rtol=2e-3)


def test_normalised_sorted_and_clamped():
    for n in np.linspace(0.1, 7.0, 57):
        m = SersicMixture.getProfile(n)
        assert m.amp.sum() == pytest.approx(1., abs=1e-12)
        assert np.all(np.diff(m.var[:, 0, 0]) >= 0)
        assert m.K <= 16
    lo = SersicMixture.getProfile(0.1)
    np.testing.assert_array_equal(lo.amp, SersicMixture.getProfile(0.29).amp)
    hi = SersicMixture.getProfile(9.0)
    np.testing.assert_array_equal(hi.amp, SersicMixture.getProfile(6.3).amp)


def test_sersic_index_and_galaxy_params():
    import tractor_amd as T
    si = SersicIndex(2.5)
    assert si.getStepSizes() == [0.01] and si.getMaxStep() == [0.1]
    assert si.getLowerBounds() == [0.29] and si.getUpperBounds() == [6.3]
    assert si.getLogPrior() == 0.
    si.setValue(7.)
    assert si.getLogPrior() == -np.inf
    g = SersicGalaxy(T.PixPos(10., 12.), T.Flux(5.), T.EllipseESoft(0.5, 0.1, -0.2),
                     SersicIndex(2.5))
    assert g.numberOfParams() == 7
    assert g.getParamNames()[-1] == 'sersicindex.SersicIndex'
    c = g.copy()
    assert c.getParams() == g.getParams() and c.getName() == 'SersicGalaxy'


def _table_profile(tab, n):
    """Reads the flat table of trh_fitter_set_sersic and evaluates it as the
    fitter does (piece: last breakpoint <= n, clamped to the first / last)."""
    pos = [0]

    def take(k):
        v = tab[pos[0]:pos[0] + k]
        pos[0] += k
        return v
    R = int(take(1)[0])
    ranges = []
    for _ in range(R):
        lo, hi, K, m = take(4)
        K, m = int(K), int(m)
        br = take(m + 1)
        comps = [(take(4 * m).reshape(m, 4), take(4 * m).reshape(m, 4)) for _ in range(K)]
        ranges.append((lo, hi, br, comps))
    m = int(take(1)[0])
    cbr, cc = take(m + 1), take(4 * m).reshape(m, 4)
    assert pos[0] == len(tab)

    def piece(br, c, x):
        i = 0
        while i + 1 < len(c) and x >= br[i + 1]:
            i += 1
        t = x - br[i]
        q = c[i]
        return ((q[0] * t + q[1]) * t + q[2]) * t + q[3]
    if n <= ranges[0][0]:
        found, n = [ranges[0]], ranges[0][0]
    elif n >= ranges[-1][1]:
        found, n = [ranges[-1]], ranges[-1][1]
    else:
        found = [r for r in ranges if r[0] <= n < r[1]]
    assert len(found) == 1          # the tabulated ranges do not overlap
    _, _, br, comps = found[0]
    amp = np.array([piece(br, a, n) for a, _ in comps])
    amp = amp / amp.sum()
    var = np.exp([piece(br, v, n) for _, v in comps])
    if n > 1.:
        core = piece(cbr, cc, n)
        amp = np.append(amp * ((1. - core) / amp.sum()), core)
        var = np.append(var, 0.)
    o = np.argsort(var, kind='stable')
    return amp[o], var[o]


def test_fitter_table_reproduces_the_mixture():
    from tractor_amd.sersic import fitter_table
    tab = fitter_table()
    nodes = [nd[0] for _, _, grid in RANGES for nd in grid]
    ns = np.concatenate([np.linspace(0.1, 7.0, 277), nodes, [lo for lo, _, _ in RANGES],
                         [1.0, 1.0 + 1e-9, 6.3]])
    for n in ns:
        a, v = _table_profile(tab, n)
        m = SersicMixture.getProfile(n)
        np.testing.assert_allclose(a, m.amp, rtol=1e-12, atol=1e-15, err_msg=str(n))
        np.testing.assert_allclose(v, m.var[:, 0, 0], rtol=1e-12, err_msg=str(n))


def test_oracle_sersic_restatement():
    """The oracle's not-a-knot CubicSpline restatement against the FITPACK
    splines of SersicMixture, and its finite-difference step rule."""
    import oracle as O
    for n in np.linspace(0.1, 7.0, 139):
        a, v = O.sersic_mixture(n)
        m = SersicMixture.getProfile(n)
        np.testing.assert_allclose(a, m.amp, rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(v, m.var[:, 0, 0], rtol=1e-12)
    # +0.01 inside a range; -0.01 where +0.01 changes the component count
    # (range ends, the core appearing above n = 1) or passes the upper bound
    assert O.sersic_fd_step(2.0) == 0.01
    assert O.sersic_fd_step(0.415) == -0.01
    assert O.sersic_fd_step(1.0) == -0.01
    assert O.sersic_fd_step(6.295) == -0.01
    assert O.sersic_fd_step(0.3) == 0.01


def test_table_matches_reference_source():
    """sersic_table.py is what tests/golden/gen_sersic_table.py writes from
    the reference's tractor/sersic.py (parsed as text; skipped where the
    reference tree is not present)."""
    import os
    src = '/root/reference/tractor/sersic.py'
    if not os.path.exists(src):
        pytest.skip('reference source not present')
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        'gen_sersic_table', os.path.join(os.path.dirname(__file__), 'golden', 'gen_sersic_table.py'))
    G = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(G)
    import tractor_amd.sersic_table as ST
    assert G.render(*G.read_tables(src)) == open(ST.__file__).read()
