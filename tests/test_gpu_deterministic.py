"""The deterministic reduction (mia_tuning.deterministic = 1) through the C ABI.

The default count adds FP64 partial sums with atomics in the order the waves
reach them, so two runs may differ in the last bits.  The deterministic mode
sorts cells in input order, gives every wave a fixed share of the work and
sums the waves' partial histograms in a fixed order (compensated); the same
count repeated, with other work in between, must give the same bits, and the
same counts and sums (to FP64 rounding) as the default path and the oracle.
"""
import numpy as np
import pytest

from measure_ia_amd import binning, hostprep, synth
from measure_ia_amd.engine import SampleArrays
from oracle import brute, prep
from paircases import compare

pytestmark = pytest.mark.gpu

DET = {"deterministic": 1}
FIELDS = ("npairs", "dd", "splus", "scross", "variance", "region_stats")


@pytest.fixture(scope="module")
def engine():
    from measure_ia_amd.engine import get_engine
    return get_engine(0)


def _clustered(scale=0.05, jk=3, seed=3):
    """A clustered box with random weights in [0.5, 1.5]; returns (data, L, host prep, sample)."""
    cat = synth.config("clustered1M", scale=scale, seed=seed)
    data = dict(cat.data)
    rng = np.random.default_rng(seed)
    data["weight"] = rng.uniform(0.5, 1.5, len(data["Position"]))
    data["weight_shape_sample"] = rng.uniform(0.5, 1.5, len(data["Position_shape_sample"]))
    p = hostprep.prepare(data)
    L = cat.boxsize
    return data, L, p, SampleArrays(p.positions, p.shapes, p.axis, p.e, p.weight, p.weight_shape, L, p.los,
                                    jk, (L / 2.0) * 2.0 / jk)


def _th(mode, L, nr=10, n2=20):
    r_bins = np.logspace(np.log10(0.1), np.log10(20.0), nr + 1)
    if mode == "rppi":
        return binning.rppi_thresholds(r_bins, np.linspace(-60, 60, n2 + 1), 0.1, 20.0)
    return binning.rmu_thresholds(r_bins, np.linspace(-1, 1, n2 + 1), 0.1, 20.0)


def _identical(a, b, tag):
    for f in FIELDS:
        x, y = getattr(a, f), getattr(b, f)
        if x is None or y is None:
            assert x is None and y is None, (tag, f)
            continue
        assert np.array_equal(x, y), (tag, f, np.abs(np.asarray(x, float) - np.asarray(y, float)).max())


def _close(a, b, tag):
    np.testing.assert_array_equal(a.npairs, b.npairs, err_msg=tag)
    for f in ("dd", "splus", "scross"):
        x, y = getattr(a, f), getattr(b, f)
        assert np.abs(x - y).max() <= 1e-12 * max(np.abs(y).max(), 1e-300), (tag, f)


@pytest.mark.parametrize("mode", ["rppi", "rmu"])
def test_repeated_counts_are_bit_identical(engine, mode):
    data, L, p, s = _clustered()
    th = _th(mode, L)
    first = engine.count(s, th, variance=True, tuning=DET)
    assert first.stats["deterministic"] == 1
    assert first.npairs[0].sum() > 100_000
    default = engine.count(s, th, variance=True)
    assert default.stats["deterministic"] == 0
    _close(first, default, "deterministic vs default")
    for k in range(3):
        # other counts in between change what the waves meet in memory and time
        engine.count(s, _th(mode, L, 7, 9))
        again = engine.count(s, th, variance=True, tuning=DET)
        _identical(again, first, f"repeat {k}")


def test_deterministic_batch_and_prepared(engine):
    cats = synth.uneven_samples(5, parent=60_000, L=205.0, lo=800, hi=20_000, shape_frac=0.3, seed=8)
    th = _th("rppi", 205.0)
    samples = []
    for c in cats:
        p = hostprep.prepare(c.data)
        samples.append(SampleArrays(p.positions, p.shapes, p.axis, p.e, p.weight, p.weight_shape, 205.0,
                                    p.los, 3, (205.0 / 2.0) * 2.0 / 3))
    a = engine.count_batch(samples, th, tuning=DET)
    assert a[0].stats["graph"] == 0 and a[0].stats["deterministic"] == 1
    ref = engine.count_batch(samples, th, tuning={"graph": -1})
    for x, y in zip(a, ref):
        _close(x, y, "batch")
    b = engine.count_batch(samples, th, tuning=DET)
    for i, (x, y) in enumerate(zip(a, b)):
        _identical(x, y, f"batch sample {i}")
    # a handle sorted in input order counts to the same bits as the plain count
    for i, smp in enumerate(samples[:2]):
        h = engine.prepare(smp, th, tuning=DET)
        one = engine.count(smp, th, tuning=DET)
        for k in range(2):
            _identical(engine.count_prepared([h], th, tuning=DET)[0], one, f"prepared {i}/{k}")
        h.release()


def test_deterministic_large_binning_in_passes(engine):
    data, L, p, s = _clustered(scale=0.02, jk=2)
    th = _th("rppi", L, 60, 60)
    a = engine.count(s, th, tuning=DET)
    assert a.stats["deterministic"] == 1
    _close(a, engine.count(s, th), "passes")
    _identical(engine.count(s, th, tuning=DET), a, "passes repeat")


def test_deterministic_against_oracle(engine):
    data, L, p, s = _clustered(scale=0.004, jk=3)
    r_bins = np.logspace(np.log10(0.1), np.log10(20.0), 11)
    sb = np.linspace(-60, 60, 21)
    th = binning.rppi_thresholds(r_bins, sb, 0.1, 20.0)
    res = engine.count(s, th, tuning=DET)
    pr = prep.prepare(data)
    jp, js = prep.jackknife_region_indices(pr["positions"], pr["shapes"], L, 3)
    ref = brute.count_pairs(pr, r_bins, sb, 0.1, 20.0, L, jk_pos=jp, jk_shape=js, num_box=27)
    compare(res, ref, p.R, 27, weighted=True)
