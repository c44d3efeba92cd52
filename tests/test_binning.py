"""Host-bisected thresholds reproduce the reference's np.log10 / np.floor bin index.

The reference formulas (measure_w_box_jk.py:142-143, 178-193;
measure_m_box_jk.py:149-150, 187-204) are evaluated directly with numpy on
random values plus every threshold and its neighbouring doubles; the
threshold rule (count of thresholds <= value) must agree everywhere, on the
benchmark binnings and on 40 random ones.
"""
import numpy as np
import pytest
from hypothesis import given, settings, strategies as st

from measure_ia_amd import binning as B
from oracle import prep

BINNINGS = [(0.1, 20.0, 10, 20, 60.0, 205.0),     # w binning of the 1M / 20k workloads
            (0.1, 60.0, 15, 20, 60.0, 205.0),     # multipoles
            (0.1, 100.0, 15, 40, 100.0, 500.0),   # large box
            (0.1, 20.0, 10, 8, None, 205.0),      # reference test configuration (pi_max = L/2)
            (1.0, 5.0, 6, 10, 5.0, 12.0)]


def _neigh(x):
    x = np.asarray(x, dtype=np.float64)
    return np.concatenate([x, B._prev(x), B._next(x)])


@pytest.mark.parametrize("cfg", BINNINGS)
def test_rppi_thresholds_match_formula(cfg):
    rmin, rmax, nr, npi, pimax, L = cfg
    rng = np.random.default_rng(0)
    r_bins, pi_bins, _ = prep.bin_edges(rmin, rmax, nr, npi, pimax, L)
    th = B.rppi_thresholds(r_bins, pi_bins, rmin, rmax)
    s = np.exp(rng.uniform(np.log(r_bins[0] ** 2 * 0.8), np.log(r_bins[-1] ** 2 * 1.2), 400_000))
    s = np.concatenate([s, _neigh(th.r2_edges), _neigh(r_bins ** 2)])
    pi = rng.uniform(pi_bins[0] * 1.2, pi_bins[-1] * 1.2, len(s))
    pi[: 3 * len(th.x2_edges)] = _neigh(th.x2_edges)
    r = np.sqrt(s)
    dom_ref = (r >= r_bins[0]) & (r < r_bins[-1]) & (pi >= pi_bins[0]) & (pi < pi_bins[-1])
    dom, ir, ip = B.bin_rppi(s, pi, th)
    np.testing.assert_array_equal(dom, dom_ref)
    irr = np.array(B.r_index(s[dom], r_bins, rmin, rmax), dtype=int)
    irr[irr >= nr] = nr - 1
    step = (pi_bins[-1] - pi_bins[0]) / npi
    ipr = np.array(np.floor(pi[dom] / step - pi_bins[0] / step), dtype=int)
    ipr[ipr >= npi] = npi - 1
    np.testing.assert_array_equal(ir[dom], irr)
    np.testing.assert_array_equal(ip[dom], ipr)


@pytest.mark.parametrize("cfg", BINNINGS)
def test_rmu_thresholds_match_formula(cfg):
    rmin, rmax, nr, nmu, pimax, L = cfg
    rng = np.random.default_rng(1)
    r_bins, _, mu_bins = prep.bin_edges(rmin, rmax, nr, nmu, pimax, L)
    th = B.rmu_thresholds(r_bins, mu_bins, rmin, rmax)
    mu = np.concatenate([rng.uniform(-1, 1, 200_000), _neigh(th.x2_edges), _neigh([-1.0, 1.0]),
                         _neigh(mu_bins)])
    _, _, im = B.bin_rmu(np.full(len(mu), th.r2_edges[0]), mu, th)
    step = 2.0 / nmu
    imr = np.array(np.floor(mu / step - mu_bins[0] / step), dtype=int)
    imr[imr >= nmu] = nmu - 1
    imr = np.where(imr < 0, imr + nmu, imr)  # numpy's negative index
    np.testing.assert_array_equal(im, imr)


def test_thresholds_are_tight():
    """Each threshold is the first double of its bin: its predecessor is in the bin below."""
    r_bins, pi_bins, _ = prep.bin_edges(0.1, 20.0, 10, 20, 60.0, 205.0)
    th = B.rppi_thresholds(r_bins, pi_bins, 0.1, 20.0)
    for k, t in enumerate(th.r2_edges[1:-1], start=1):
        assert B.r_index(t, r_bins, 0.1, 20.0) >= k
        assert B.r_index(B._prev(t), r_bins, 0.1, 20.0) < k
    assert np.sqrt(th.r2_edges[0]) >= r_bins[0] and np.sqrt(B._prev(th.r2_edges[0])) < r_bins[0]
    assert np.sqrt(th.r2_edges[-1]) >= r_bins[-1] and np.sqrt(B._prev(th.r2_edges[-1])) < r_bins[-1]


def test_ordered_keys_roundtrip():
    x = np.array([-1e300, -2.5, -0.0, 0.0, 1e-300, 3.0, 1e300])
    k = B.to_ordered(x)
    assert np.all(np.diff(k) >= 0)
    np.testing.assert_array_equal(np.abs(B.from_ordered(k)), np.abs(x))


def test_planted_log10_slip_is_caught():
    """A threshold shifted by one ulp must disagree with the formula on its neighbour."""
    r_bins, pi_bins, _ = prep.bin_edges(0.1, 20.0, 10, 20, 60.0, 205.0)
    th = B.rppi_thresholds(r_bins, pi_bins, 0.1, 20.0)
    bad = th.r2_edges.copy()
    bad[3] = B._next(bad[3])
    s = th.r2_edges[3]
    assert np.searchsorted(bad[1:-1], s, side="right") != np.searchsorted(th.r2_edges[1:-1], s,
                                                                          side="right")



@settings(max_examples=40, deadline=None, derandomize=True)
@given(rmin=st.floats(0.01, 5.0), ratio=st.floats(1.5, 1000.0), nr=st.integers(1, 60),
       n2=st.integers(1, 60), pimax=st.floats(0.5, 300.0))
def test_random_binnings_match_formula(rmin, ratio, nr, n2, pimax):
    """Random binnings (the sizes a user may choose): the threshold rule equals
    the reference's index formulas at random values and around every threshold,
    for (r_p, pi) and for mu."""
    rmax = rmin * ratio
    r_bins, pi_bins, mu_bins = prep.bin_edges(rmin, rmax, nr, n2, pimax, 2 * pimax)
    rng = np.random.default_rng(nr * 1000 + n2)
    th = B.rppi_thresholds(r_bins, pi_bins, rmin, rmax)
    s = np.exp(rng.uniform(np.log(r_bins[0] ** 2 * 0.9), np.log(r_bins[-1] ** 2 * 1.1), 20_000))
    s = np.concatenate([s, _neigh(th.r2_edges), _neigh(r_bins ** 2)])
    pi = rng.uniform(pi_bins[0] * 1.1, pi_bins[-1] * 1.1, len(s))
    pi[: 3 * len(th.x2_edges)] = _neigh(th.x2_edges)
    r = np.sqrt(s)
    dom, ir, ip = B.bin_rppi(s, pi, th)
    np.testing.assert_array_equal(dom, (r >= r_bins[0]) & (r < r_bins[-1]) & (pi >= pi_bins[0])
                                  & (pi < pi_bins[-1]))
    irr = np.array(B.r_index(s[dom], r_bins, rmin, rmax), dtype=int)
    irr[irr >= nr] = nr - 1
    step = (pi_bins[-1] - pi_bins[0]) / n2
    ipr = np.array(np.floor(pi[dom] / step - pi_bins[0] / step), dtype=int)
    ipr[ipr >= n2] = n2 - 1
    np.testing.assert_array_equal(ir[dom], irr)
    np.testing.assert_array_equal(ip[dom], ipr)
    thm = B.rmu_thresholds(r_bins, mu_bins, rmin, rmax)
    mu = np.concatenate([rng.uniform(-1, 1, 20_000), _neigh(thm.x2_edges), _neigh([-1.0, 1.0])])
    _, _, im = B.bin_rmu(np.full(len(mu), thm.r2_edges[0]), mu, thm)
    mstep = 2.0 / n2
    imr = np.array(np.floor(mu / mstep - mu_bins[0] / mstep), dtype=int)
    imr[imr >= n2] = n2 - 1
    imr = np.where(imr < 0, imr + n2, imr)
    np.testing.assert_array_equal(im, imr)
