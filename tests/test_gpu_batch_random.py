"""Randomised batches with dyadic weights: the batched count equals every
sample's own count bit for bit.

Each seeded case draws 2-6 samples of uneven size (200-5000 positions cut
from one clustered box, 10-60 % of them shapes), a jackknife split per sample
(none, 2 or 3 per side), a line of sight per sample and (r_p, pi) or (r, mu)
binning.  Weights come from {0.5, 1, 2}, so every DD sum is exact in FP64
whatever the order of the adds.  The batch is counted three times (captured
in a hipGraph, replayed from it, and with plain launches) and each sample
once on its own: pair counts, DD and every DD tally must agree exactly, S+D
to FP64 rounding.
Every third case also checks its first sample against the brute-force
oracle (measure_w_box_jk.py:142-211 restated in oracle/brute.py).
"""
import numpy as np
import pytest

from measure_ia_amd import binning, hostprep
from measure_ia_amd.engine import SampleArrays
from measure_ia_amd.synth import uneven_samples
from oracle import brute, prep
from paircases import compare, dyadic_weights

pytestmark = pytest.mark.gpu

L = 60.0
R_MIN, R_MAX, NR, NPI, PI_MAX = 0.1, 10.0, 10, 8, 20.0


@pytest.fixture(scope="module")
def engine():
    from measure_ia_amd.engine import get_engine
    return get_engine(0)


def _case(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 7))
    cats = uneven_samples(n_samples=n, parent=20000, L=L, lo=200, hi=5000,
                          shape_frac=float(rng.uniform(0.1, 0.6)), seed=seed)
    out = []
    for cat in cats:
        d = dict(cat.data)
        d["LOS"] = int(rng.integers(0, 3))
        d["weight"] = dyadic_weights(rng, len(d["Position"]))
        d["weight_shape_sample"] = dyadic_weights(rng, len(d["Position_shape_sample"]))
        out.append((d, int(rng.choice([0, 2, 3]))))
    mode = "rppi" if rng.uniform() < 0.5 else "rmu"
    return out, mode


def _engine_sample(p, jk):
    return SampleArrays(positions=p.positions, shapes=p.shapes, axis=p.axis, e=p.e, weight=p.weight,
                        weight_shape=p.weight_shape, boxsize=L, los=p.los, jk_per_side=jk,
                        jk_sub=(L / 2.0) * 2.0 / jk if jk else 0.0)


def _same(a, b, what):
    np.testing.assert_array_equal(a.npairs, b.npairs, err_msg=what)
    np.testing.assert_array_equal(a.dd, b.dd, err_msg=what)
    scale = max(np.abs(b.splus).max(), 1e-300)
    assert np.abs(a.splus - b.splus).max() <= 1e-12 * scale, what


@pytest.mark.parametrize("seed", range(900, 912))
def test_random_batch_equals_single_counts(engine, seed):
    samples, mode = _case(seed)
    r_bins, pi_bins, mu_bins = prep.bin_edges(R_MIN, R_MAX, NR, NPI, PI_MAX, L)
    if mode == "rppi":
        th = binning.rppi_thresholds(r_bins, pi_bins, R_MIN, R_MAX)
        sb = pi_bins
    else:
        th = binning.rmu_thresholds(r_bins, mu_bins, R_MIN, R_MAX)
        sb = mu_bins
    preps = [hostprep.prepare(d) for d, _ in samples]
    es = [_engine_sample(p, jk) for p, (_, jk) in zip(preps, samples)]
    print("seed", seed, mode, [(len(p.positions), len(p.shapes), jk, p.los) for p, (_, jk) in
                               zip(preps, samples)], flush=True)
    graph = engine.count_batch(es, th)  # captured and launched
    replay = engine.count_batch(es, th)  # the same key: replayed
    plain = engine.count_batch(es, th, tuning={"graph": -1})
    for i, (s, g, r, q) in enumerate(zip(es, graph, replay, plain)):
        one = engine.count(s, th, method="fast")
        _same(g, one, f"sample {i}: captured graph vs single count")
        _same(r, one, f"sample {i}: graph replay vs single count")
        _same(q, one, f"sample {i}: plain launches vs single count")
    if seed % 3 == 0:
        (d, jk), p = samples[0], preps[0]
        pr = prep.prepare(d)
        K = jk ** 3 if jk else 0
        jp = js = None
        if K:
            jp, js = prep.jackknife_region_indices(pr["positions"], pr["shapes"], L, jk)
        ref = brute.count_pairs(pr, r_bins, sb, R_MIN, R_MAX, L, mode=mode, jk_pos=jp, jk_shape=js,
                                num_box=K)
        np.testing.assert_array_equal(graph[0].dd[0], ref["DD"])
        if K:
            np.testing.assert_array_equal(graph[0].dd[1:], ref["DD_jk"])
        compare(graph[0], ref, p.R, K, weighted=True)
