"""Batched counting (mia_count_batch) against the oracle and single counts.

Samples of uneven size cut from one clustered box, with weights, counted in
one persistent launch: every sample's integer pair counts and jackknife
tallies must equal the oracle's bit for bit, weighted DD and the S+D / SxD sums
within 1e-6, and each sample must equal its own single-sample count (integer
counts exactly, FP64 sums up to atomic-order rounding).  The batch also mixes
jackknife settings, line-of-sight axes and an empty shape sample.
"""
import numpy as np
import pytest

from measure_ia_amd import binning, hostprep
from measure_ia_amd.engine import SampleArrays
from measure_ia_amd.synth import uneven_samples
from oracle import brute, prep
from paircases import compare

pytestmark = pytest.mark.gpu

L = 60.0
R_MIN, R_MAX, NR, NPI, PI_MAX = 0.1, 10.0, 10, 8, 20.0


@pytest.fixture(scope="module")
def engine():
    from measure_ia_amd.engine import get_engine
    return get_engine(0)


def _samples():
    cats = uneven_samples(n_samples=5, parent=20000, L=L, lo=300, hi=6000, seed=21)
    specs = [(3, 2), (3, 2), (2, 2), (0, 2), (3, 0)]  # (jk per side, LOS) per sample
    out = []
    for cat, (jk, los) in zip(cats, specs):
        d = dict(cat.data)
        d["LOS"] = los
        out.append((d, jk))
    return out


def _engine_sample(p, jk, weighted=True):
    return SampleArrays(positions=p.positions, shapes=p.shapes, axis=p.axis, e=p.e,
                        weight=p.weight if weighted else None,
                        weight_shape=p.weight_shape if weighted else None,
                        boxsize=L, los=p.los, jk_per_side=jk, jk_sub=(L / 2.0) * 2.0 / jk if jk else 0.0)


@pytest.mark.parametrize("mode", ["rppi", "rmu"])
def test_batch_matches_oracle_and_single(engine, mode):
    r_bins, pi_bins, mu_bins = prep.bin_edges(R_MIN, R_MAX, NR, NPI, PI_MAX, L)
    if mode == "rppi":
        th = binning.rppi_thresholds(r_bins, pi_bins, R_MIN, R_MAX)
        sb = pi_bins
    else:
        th = binning.rmu_thresholds(r_bins, mu_bins, R_MIN, R_MAX)
        sb = mu_bins
    samples = _samples()
    preps = [hostprep.prepare(d) for d, _ in samples]
    es = [_engine_sample(p, jk) for p, (_, jk) in zip(preps, samples)]
    print("sizes", [(len(p.positions), len(p.shapes), jk) for p, (_, jk) in zip(preps, samples)],
          flush=True)
    batch = engine.count_batch(es, th)
    assert len(batch) == len(es)
    for i, ((d, jk), p, s, b) in enumerate(zip(samples, preps, es, batch)):
        pr = prep.prepare(d)
        K = jk ** 3 if jk else 0
        jp = js = None
        if K:
            jp, js = prep.jackknife_region_indices(pr["positions"], pr["shapes"], L, jk)
        ref = brute.count_pairs(pr, r_bins, sb, R_MIN, R_MAX, L, mode=mode, jk_pos=jp, jk_shape=js,
                                num_box=K)
        errs = compare(b, ref, p.R, K, weighted=True)
        one = engine.count(s, th, method="fast")
        np.testing.assert_array_equal(b.npairs, one.npairs)
        for name in ("dd", "splus"):
            x, y = getattr(b, name), getattr(one, name)
            assert np.abs(x - y).max() <= 1e-12 * max(np.abs(y).max(), 1e-300), (i, name)
        print(i, mode, errs, flush=True)


def test_batch_with_empty_shape_sample(engine):
    r_bins, pi_bins, _ = prep.bin_edges(R_MIN, R_MAX, NR, NPI, PI_MAX, L)
    th = binning.rppi_thresholds(r_bins, pi_bins, R_MIN, R_MAX)
    (d, jk), = _samples()[:1]
    p = hostprep.prepare(d)
    full = _engine_sample(p, jk)
    empty = SampleArrays(positions=p.positions, shapes=np.zeros((0, 3)), axis=np.zeros((0, 2)),
                         e=np.zeros(0), weight=p.weight, weight_shape=np.zeros(0), boxsize=L,
                         los=p.los, jk_per_side=jk, jk_sub=full.jk_sub)
    a, b = engine.count_batch([empty, full], th)
    assert a.npairs.sum() == 0 and np.all(a.splus == 0) and np.all(a.dd == 0)
    one = engine.count(full, th)
    np.testing.assert_array_equal(b.npairs, one.npairs)


@pytest.mark.parametrize("mode", ["rppi", "rmu"])
def test_batch_dyadic_weights_exact(engine, mode):
    """Weights from {0.5, 1, 2}: every DD sum is exact in FP64, so the batched
    count, the single count and the oracle agree bit for bit on DD and on every
    DD tally (measure_w_box_jk.py:207-211), whatever the atomic order.  A lost,
    repeated or misattributed weighted add fails this deterministically."""
    from paircases import dyadic_weights
    rng = np.random.default_rng(77)
    r_bins, pi_bins, mu_bins = prep.bin_edges(R_MIN, R_MAX, NR, NPI, PI_MAX, L)
    if mode == "rppi":
        th = binning.rppi_thresholds(r_bins, pi_bins, R_MIN, R_MAX)
        sb = pi_bins
    else:
        th = binning.rmu_thresholds(r_bins, mu_bins, R_MIN, R_MAX)
        sb = mu_bins
    samples = _samples()
    for d, _ in samples:
        d["weight"] = dyadic_weights(rng, len(d["Position"]))
        d["weight_shape_sample"] = dyadic_weights(rng, len(d["Position_shape_sample"]))
    preps = [hostprep.prepare(d) for d, _ in samples]
    es = [_engine_sample(p, jk) for p, (_, jk) in zip(preps, samples)]
    batch = engine.count_batch(es, th)
    singles = [engine.count(s, th, method="fast") for s in es]
    for i, ((d, jk), p, b, one) in enumerate(zip(samples, preps, batch, singles)):
        pr = prep.prepare(d)
        K = jk ** 3 if jk else 0
        jp = js = None
        if K:
            jp, js = prep.jackknife_region_indices(pr["positions"], pr["shapes"], L, jk)
        ref = brute.count_pairs(pr, r_bins, sb, R_MIN, R_MAX, L, mode=mode, jk_pos=jp, jk_shape=js,
                                num_box=K)
        print("sample", i, len(pr["positions"]), len(pr["shapes"]), jk, flush=True)
        np.testing.assert_array_equal(b.npairs, one.npairs)
        np.testing.assert_array_equal(b.dd, one.dd)
        np.testing.assert_array_equal(b.dd[0], ref["DD"])
        if K:
            np.testing.assert_array_equal(b.dd[1:], ref["DD_jk"])
        compare(b, ref, p.R, K, weighted=True)
        compare(one, ref, p.R, K, weighted=True)
