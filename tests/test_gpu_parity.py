"""HIP path against the FP64 oracle on seeded inputs (through the C ABI).

DD and every jackknife tally are compared bit-exactly (integer pair counts,
and unit-weight DD); S+D, SxD and the S+D tallies within 1e-6 of
max(|ref|, sum|terms|) per bin; weighted DD within 1e-6 relative.
"""
import functools

import numpy as np
import pytest

from paircases import CASES, compare, engine_sample, make_case, oracle_counts, thresholds

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    from measure_ia_amd.engine import get_engine
    return get_engine(0)


@functools.lru_cache(maxsize=None)
def _oracle(case, mode):
    data, L, cfg = make_case(case)
    return oracle_counts(data, L, cfg, mode)


@pytest.mark.parametrize("method", ["plain", "fast"])
@pytest.mark.parametrize("mode", ["rppi", "rmu"])
@pytest.mark.parametrize("case", CASES)
def test_parity(engine, case, mode, method):
    data, L, cfg = make_case(case)
    pr, ref, _ = _oracle(case, mode)
    p, sample = engine_sample(data, L, cfg)
    assert p.R == pr["R"]
    th = thresholds(cfg, L, mode)
    res = engine.count(sample, th, method=method, variance=True)
    K = cfg["L_sub"] ** 3 if cfg["L_sub"] else 0
    errs = compare(res, ref, p.R, K, cfg["weights"])
    v = res.variance / (2 * p.R) ** 2
    assert np.all(np.abs(v - ref["variance"]) <= 1e-6 * np.maximum(ref["variance"], 1e-300))
    print(case, mode, method, errs, res.stats)


def test_region_ids(engine):
    from oracle import prep
    rng = np.random.default_rng(3)
    L = 9.0
    x = rng.uniform(0, L, (5000, 3))
    x[:300] = np.round(x[:300])  # many objects on faces
    x[300:310] = 0.0
    x[310:320] = L
    for n in (1, 2, 3, 4):
        ref, _ = prep.jackknife_region_indices(x, x[:1], L, n)
        got = engine.region_ids(x, n, (L / 2.0) * 2.0 / n)
        np.testing.assert_array_equal(got, ref)


def test_fast_equals_plain_determinism(engine):
    data, L, cfg = make_case("clustered_jk8")
    p, sample = engine_sample(data, L, cfg)
    th = thresholds(cfg, L, "rppi")
    a = engine.count(sample, th, method="fast")
    b = engine.count(sample, th, method="fast")
    c = engine.count(sample, th, method="plain")
    # integer counts are identical between runs and between the two methods
    np.testing.assert_array_equal(a.npairs, b.npairs)
    np.testing.assert_array_equal(a.npairs, c.npairs)
    # FP64 sums only move by the rounding of the (unordered) atomic additions
    scale = np.abs(a.splus).max()
    assert np.abs(a.splus - b.splus).max() <= 1e-12 * scale
    assert np.abs(a.splus - c.splus).max() <= 1e-12 * scale
    assert np.abs(a.scross - c.scross).max() <= 1e-12 * np.abs(a.scross).max()
