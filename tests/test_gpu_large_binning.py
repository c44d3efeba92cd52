"""Binnings larger than the default LDS budget.

3,600 (r_p, pi) bins no longer fit 4 wave histograms per workgroup, so the
batched kernel runs 2 waves per workgroup; 10,000 bins do not fit even one,
so mia_count counts them with the plain kernel (reported in stats.method) and
the batched entry point refuses them.  Both agree with the oracle.
"""
import numpy as np
import pytest

from measure_ia_amd import binning, hostprep
from measure_ia_amd._capi import MiaError
from measure_ia_amd.engine import SampleArrays
from oracle import brute, prep

pytestmark = pytest.mark.gpu

L = 40.0


@pytest.fixture(scope="module")
def engine():
    from measure_ia_amd.engine import get_engine
    return get_engine(0)


def _case(nr, npi):
    rng = np.random.default_rng(nr)
    pos = rng.uniform(0, L, (1500, 3))
    shp = rng.uniform(0, L, (600, 3))
    th_ax = rng.uniform(0, np.pi, 600)
    d = {"Position": pos, "Position_shape_sample": shp, "LOS": 2, "q": rng.uniform(0.4, 1.0, 600),
         "Axis_Direction": np.stack([np.cos(th_ax), np.sin(th_ax)], 1)}
    r_bins, pi_bins, _ = prep.bin_edges(0.2, 10.0, nr, npi, 12.0, L)
    th = binning.rppi_thresholds(r_bins, pi_bins, 0.2, 10.0)
    p = hostprep.prepare(d)
    s = SampleArrays(p.positions, p.shapes, p.axis, p.e, None, None, L, 2, 2, L / 2.0)
    pr = prep.prepare(d)
    jp, js = prep.jackknife_region_indices(pr["positions"], pr["shapes"], L, 2)
    ref = brute.count_pairs(pr, r_bins, pi_bins, 0.2, 10.0, L, jk_pos=jp, jk_shape=js, num_box=8)
    return s, th, ref, p.R


@pytest.mark.parametrize("nr, npi, method_used", [(60, 60, 1), (100, 100, 0)])
def test_large_binning(engine, nr, npi, method_used):
    s, th, ref, R = _case(nr, npi)
    res = engine.count(s, th, method="fast")
    assert res.stats["method"] == method_used
    np.testing.assert_array_equal(res.npairs[0], ref["NP"])
    np.testing.assert_array_equal(res.npairs[1:], ref["NP_jk"])
    err = np.abs(res.splus[0] / (2 * R) - ref["SplusD"]) / np.maximum(np.abs(ref["SplusD"]),
                                                                      ref["SplusD_abs"] + 1e-300)
    assert err.max() <= 1e-6
    if method_used == 0:
        with pytest.raises(MiaError) as e:
            engine.count_batch([s], th)
        assert e.value.code == -5
