"""mia_count_noalloc -- the allocation-free variant of mia_count (MIA_FAST) that
carves every buffer from the library's static device workspace -- against
mia_count on the same inputs, through the C ABI.

Two calls in a row share the workspace, the second with a smaller sample, a
different jackknife split and line of sight and no weights; each must equal
mia_count: integer pair counts, tallies and unit-weight DD bit for bit,
weighted DD and the S+D / SxD sums to 1e-12 of the bin scale (atomic order),
region statistics (counts exactly).  A sample too large for the workspace
is refused with MIA_E_UNSUPPORTED before anything is launched.
"""
import ctypes

import numpy as np
import pytest

from measure_ia_amd import _capi, binning, hostprep
from measure_ia_amd.engine import SampleArrays
from measure_ia_amd.synth import clustered
from oracle import prep

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    from measure_ia_amd.engine import get_engine
    return get_engine(0)


def _noalloc(eng, s, th, variance=False):
    ms, keep, _ = eng._prepare(s)
    K = s.jk_per_side ** 3 if s.jk_per_side > 0 else 0
    bins, keep_b = eng._bins(th, variance)
    layout, buf, c = eng._outputs(K, th.n_r * th.n_2, variance)
    stream = eng.torch.cuda.current_stream(eng.dev).cuda_stream
    rc = eng.lib.mia_count_noalloc(eng.ctx, ctypes.byref(ms), ctypes.byref(bins), ctypes.byref(c), None,
                                   ctypes.c_void_p(stream))
    _capi.check(eng.lib, eng.ctx, rc)
    return eng._host((layout, buf), K, th, {})


def _sample(n_pos, n_shape, L, jk, los, weighted, seed):
    cat = clustered(n_pos, n_shape, L, max(1, n_pos // 400), 1.0, 40, 0.2, 0.1, seed=seed)
    d = dict(cat.data)
    d["LOS"] = los
    if weighted:
        rng = np.random.default_rng(seed)
        d["weight"] = rng.uniform(0.5, 1.5, n_pos)
        d["weight_shape_sample"] = rng.uniform(0.5, 1.5, n_shape)
    p = hostprep.prepare(d)
    return SampleArrays(positions=p.positions, shapes=p.shapes, axis=p.axis, e=p.e,
                        weight=p.weight if weighted else None,
                        weight_shape=p.weight_shape if weighted else None, boxsize=L, los=p.los,
                        jk_per_side=jk, jk_sub=(L / 2.0) * 2.0 / jk if jk else 0.0)


def _equal(a, b, weighted):
    np.testing.assert_array_equal(a.npairs, b.npairs)
    if weighted:
        assert np.abs(a.dd - b.dd).max() <= 1e-12 * np.abs(b.dd).max()
    else:
        np.testing.assert_array_equal(a.dd, b.dd)
    for name in ("splus", "scross"):
        x, y = getattr(a, name), getattr(b, name)
        assert np.abs(x - y).max() <= 1e-12 * max(np.abs(y).max(), 1e-300), name
    if b.region_stats is not None:
        np.testing.assert_array_equal(a.region_stats[:, :2], b.region_stats[:, :2])
        np.testing.assert_allclose(a.region_stats[:, 2:], b.region_stats[:, 2:], rtol=1e-13)


def test_noalloc_equals_mia_count_twice(engine):
    L = 120.0
    r_bins, pi_bins, mu_bins = prep.bin_edges(0.1, 20.0, 10, 20, 60.0, L)
    th = binning.rppi_thresholds(r_bins, pi_bins, 0.1, 20.0)
    big = _sample(200_000, 60_000, L, 3, 2, True, seed=5)
    small = _sample(20_000, 6_000, L, 2, 0, False, seed=6)
    print("workspace bytes", engine.lib.mia_noalloc_workspace_bytes(), flush=True)
    a = _noalloc(engine, big, th)
    _equal(a, engine.count(big, th, method="fast"), True)
    assert a.npairs[0].sum() > 1_000_000
    b = _noalloc(engine, small, th)
    _equal(b, engine.count(small, th, method="fast"), False)
    assert b.npairs[0].sum() > 10_000
    # (r, mu) through the same workspace
    thm = binning.rmu_thresholds(r_bins, mu_bins, 0.1, 20.0)
    _equal(_noalloc(engine, small, thm), engine.count(small, thm, method="fast"), False)


def test_noalloc_refuses_a_sample_beyond_the_workspace(engine):
    import torch
    L = 205.0
    n = 8_000_000
    dev = torch.device("cuda", 0)
    pos = torch.rand((n, 3), dtype=torch.float64, device=dev) * L
    shp = torch.rand((1, 3), dtype=torch.float64, device=dev) * L
    s = SampleArrays(positions=pos, shapes=shp, axis=torch.tensor([[1.0, 0.0]], dtype=torch.float64, device=dev),
                     e=torch.zeros(1, dtype=torch.float64, device=dev), boxsize=L, los=2, jk_per_side=0)
    r_bins, pi_bins, _ = prep.bin_edges(0.1, 20.0, 10, 20, 60.0, L)
    th = binning.rppi_thresholds(r_bins, pi_bins, 0.1, 20.0)
    with pytest.raises(_capi.MiaError) as ei:
        _noalloc(engine, s, th)
    assert ei.value.code == -5 and "static workspace" in str(ei.value)
