"""Binnings larger than the default LDS budget.

3,600 and 10,000 (r_p, pi) bins leave too few waves per CU for one pass, so
the batched kernel counts them in passes over r-bin ranges (2 and 5 passes
by default); the counts agree with the oracle, with a
single-sample batch, and with explicit pass sizes of 1, 7 and all r bins.  A thin shell (r in
[10, 10.01) over 10 log bins) is binned by scanning every FP64 threshold,
since the FP32 guess is too coarse there.
"""
import numpy as np
import pytest

from measure_ia_amd import binning, hostprep
from measure_ia_amd.engine import SampleArrays
from oracle import brute, prep

pytestmark = pytest.mark.gpu

L = 40.0


@pytest.fixture(scope="module")
def engine():
    from measure_ia_amd.engine import get_engine
    return get_engine(0)


def _case(nr, npi, weighted=False):
    rng = np.random.default_rng(nr)
    pos = rng.uniform(0, L, (1500, 3))
    shp = rng.uniform(0, L, (600, 3))
    th_ax = rng.uniform(0, np.pi, 600)
    d = {"Position": pos, "Position_shape_sample": shp, "LOS": 2, "q": rng.uniform(0.4, 1.0, 600),
         "Axis_Direction": np.stack([np.cos(th_ax), np.sin(th_ax)], 1)}
    if weighted:
        d["weight"] = rng.uniform(0.5, 1.5, len(pos))
        d["weight_shape_sample"] = rng.uniform(0.5, 1.5, len(shp))
    r_bins, pi_bins, _ = prep.bin_edges(0.2, 10.0, nr, npi, 12.0, L)
    th = binning.rppi_thresholds(r_bins, pi_bins, 0.2, 10.0)
    p = hostprep.prepare(d)
    s = SampleArrays(p.positions, p.shapes, p.axis, p.e, p.weight if weighted else None,
                     p.weight_shape if weighted else None, L, 2, 2, L / 2.0)
    pr = prep.prepare(d)
    jp, js = prep.jackknife_region_indices(pr["positions"], pr["shapes"], L, 2)
    ref = brute.count_pairs(pr, r_bins, pi_bins, 0.2, 10.0, L, jk_pos=jp, jk_shape=js, num_box=8)
    return s, th, ref, p.R


@pytest.mark.parametrize("nr, npi, weighted", [(60, 60, False), (100, 100, False), (100, 100, True)])
def test_large_binning(engine, nr, npi, weighted):
    s, th, ref, R = _case(nr, npi, weighted)
    res = engine.count(s, th, method="fast", variance=True)
    assert res.stats["method"] == 1
    np.testing.assert_array_equal(res.npairs[0], ref["NP"])
    np.testing.assert_array_equal(res.npairs[1:], ref["NP_jk"])
    err = np.abs(res.splus[0] / (2 * R) - ref["SplusD"]) / np.maximum(np.abs(ref["SplusD"]),
                                                                      ref["SplusD_abs"] + 1e-300)
    assert err.max() <= 1e-6
    err = np.abs(res.scross / (2 * R) - ref["ScrossD"]) / np.maximum(np.abs(ref["ScrossD"]),
                                                                     ref["ScrossD_abs"] + 1e-300)
    assert err.max() <= 1e-6
    if weighted:
        assert np.abs(res.dd[0] - ref["DD"]).max() <= 1e-12 * np.abs(ref["DD"]).max()
    v = res.variance / (2 * R) ** 2
    assert np.all(np.abs(v - ref["variance"]) <= 1e-6 * np.maximum(ref["variance"], 1e-300))
    # a one-sample batch and explicit pass sizes give the same counts
    b = engine.count_batch([s], th)[0]
    np.testing.assert_array_equal(b.npairs, res.npairs)
    for per_pass in (1, 7, nr):
        t = engine.count(s, th, method="fast", tuning={"r_bins_per_pass": per_pass})
        np.testing.assert_array_equal(t.npairs, res.npairs)
        assert np.abs(t.splus - res.splus).max() <= 1e-12 * np.abs(res.splus).max()


@pytest.mark.parametrize("mode", ["rppi", "rmu"])
def test_thin_shell_scans_every_threshold(engine, mode):
    """r in [10, 10.01) over 10 log bins: the FP32 bin guess is no longer
    within 5% of a bin, so every in-domain pair takes the exact FP64 path and
    counts every threshold (stats.rechecks == pairs); counts equal the oracle."""
    rng = np.random.default_rng(77)
    pos = rng.uniform(0, L, (12000, 3))
    shp = rng.uniform(0, L, (3000, 3))
    th_ax = rng.uniform(0, np.pi, 3000)
    d = {"Position": pos, "Position_shape_sample": shp, "LOS": 2, "q": rng.uniform(0.4, 1.0, 3000),
         "Axis_Direction": np.stack([np.cos(th_ax), np.sin(th_ax)], 1)}
    r_bins, pi_bins, mu_bins = prep.bin_edges(10.0, 10.01, 10, 8, 12.0, L)
    second = pi_bins if mode == "rppi" else mu_bins
    th = (binning.rppi_thresholds if mode == "rppi" else binning.rmu_thresholds)(r_bins, second, 10.0, 10.01)
    p = hostprep.prepare(d)
    s = SampleArrays(p.positions, p.shapes, p.axis, p.e, None, None, L, 2, 2, L / 2.0)
    pr = prep.prepare(d)
    jp, js = prep.jackknife_region_indices(pr["positions"], pr["shapes"], L, 2)
    ref = brute.count_pairs(pr, r_bins, second, 10.0, 10.01, L, mode=mode, jk_pos=jp, jk_shape=js, num_box=8)
    res = engine.count(s, th, method="fast")
    assert res.stats["method"] == 1
    assert ref["NP"].sum() > 2000
    np.testing.assert_array_equal(res.npairs[0], ref["NP"])
    np.testing.assert_array_equal(res.npairs[1:], ref["NP_jk"])
    assert res.stats["rechecks"] == ref["NP"].sum()
    err = np.abs(res.splus[0] / (2 * p.R) - ref["SplusD"]) / np.maximum(np.abs(ref["SplusD"]),
                                                                        ref["SplusD_abs"] + 1e-300)
    assert err.max() <= 1e-6
