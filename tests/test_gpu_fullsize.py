"""Parity at the benchmark sizes.

* 20,000 uniform galaxies (shapes = positions), w_gg + w_g+ binning, 27
  regions: the whole count against the oracle's KDTree counter (the
  reference's tree path, oracle/tree.py) -- every pair count and jackknife
  tally equal, S+D, SxD within 1e-6 of max(|ref|, sum|terms|).
* 1,000,000 clustered positions / 200,000 shapes, w_gg + w_g+ binning, 27
  regions: the whole count against the oracle's KDTree counter over all 200
  shape chunks (about a minute on 16 cores) -- pair counts and jackknife
  tallies equal, S+D, SxD and the projected w_g+ per r_p bin within 1e-6.
* The 2M / 500k multipole box: 16 seeded chunks of 10 shapes against all 2M
  positions in (r, mu), against the oracle's KDTree counter.
* Both boxes also through identities the counts must satisfy exactly:
    - leave one region out: a fresh count of the sample without the objects
      of region i equals total - tally[i] (measure_w_box_jk.py:258-287 builds
      realisation i the same way);
    - line-of-sight permutation: the same box with its columns permuted and
      LOS moved with them gives the same counts;
    - batched launch: 12 uneven weighted samples counted in one launch equal
      12 single counts.
* The large box (20,000,000 clustered positions, L = 500, 15 log r_p bins
  over 0.1-100, 40 pi bins over |pi| < 100, 125 regions): 160 seeded shapes
  (16 chunks of 10) against all 20M positions, about 1.6e8 pairs, against the
  oracle's KDTree counter -- the kernel at the large box's own density,
  binning and region count, checked pair for pair.
"""
import numpy as np
import pytest

from measure_ia_amd import binning, hostprep, synth
from measure_ia_amd.engine import SampleArrays

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    from measure_ia_amd.engine import get_engine
    return get_engine(0)


def _w_thresholds(L):
    r_bins = np.logspace(np.log10(0.1), np.log10(20.0), 11)
    return binning.rppi_thresholds(r_bins, np.linspace(-60, 60, 21), 0.1, 20.0), r_bins


def _sample(p, L, jk, sel_p=None, sel_s=None, weighted=False):
    sp = slice(None) if sel_p is None else sel_p
    ss = slice(None) if sel_s is None else sel_s
    return SampleArrays(p.positions[sp], p.shapes[ss], p.axis[ss], p.e[ss],
                        p.weight[sp] if weighted else None, p.weight_shape[ss] if weighted else None,
                        L, p.los, jk, (L / 2.0) * 2.0 / jk if jk else 0.0)


def _rel(got, ref, scale):
    den = np.maximum(np.abs(ref), scale)
    return np.where(den > 0, np.abs(got - ref) / np.where(den > 0, den, 1), np.abs(got - ref))


def test_uniform20k_against_tree_oracle(engine):
    from oracle import prep, tree
    cat = synth.config("uniform20k")
    L = cat.boxsize
    th, r_bins = _w_thresholds(L)
    p = hostprep.prepare(cat.data)
    res = engine.count(_sample(p, L, 3), th)
    pr = prep.prepare(cat.data)
    jp, js = prep.jackknife_region_indices(pr["positions"], pr["shapes"], L, 3)
    ref = tree.count_pairs(pr, r_bins, np.linspace(-60, 60, 21), 0.1, 20.0, L, jk_pos=jp, jk_shape=js,
                           num_box=27, num_nodes=min(16, tree.default_workers()), abs_sums=True)
    print("pairs", ref["npairs"], "oracle s", round(ref["seconds"], 1), flush=True)
    assert res.npairs[0].sum() == ref["npairs"] > 6_000_000
    np.testing.assert_array_equal(res.dd[0], ref["DD"])
    np.testing.assert_array_equal(res.dd[1:], ref["DD_jk"])
    R = p.R
    assert _rel(res.splus[0] / (2 * R), ref["SplusD"], ref["SplusD_abs"]).max() <= 1e-6
    assert _rel(res.scross / (2 * R), ref["ScrossD"], ref["ScrossD_abs"]).max() <= 1e-6
    assert _rel(res.splus[1:], ref["SplusD_jk"], ref["SplusD_jk_abs"]).max() <= 1e-6
    # Few persistent waves make every wave's share large enough that the
    # waves also claim tail chunks from the counter: the same counts.
    for waves in (16, 333):
        t = engine.count(_sample(p, L, 3), th, tuning={"waves": waves})
        assert t.stats["waves"] == waves
        np.testing.assert_array_equal(t.npairs, res.npairs)
        np.testing.assert_array_equal(t.dd, res.dd)
        assert np.abs(t.splus - res.splus).max() <= 1e-12 * np.abs(res.splus).max()


@pytest.fixture(scope="module")
def clustered():
    cat = synth.config("clustered1M")
    return cat, hostprep.prepare(cat.data)


def test_clustered1M_leave_one_region_out(engine, clustered):
    cat, p = clustered
    L = cat.boxsize
    th, _ = _w_thresholds(L)
    full = engine.count(_sample(p, L, 3), th)
    assert full.npairs[0].sum() > 3_000_000_000
    jk_sub = (L / 2.0) * 2.0 / 3
    reg_p = engine.region_ids(p.positions, 3, jk_sub)
    reg_s = engine.region_ids(p.shapes, 3, jk_sub)
    for i in (0, 13, 26):
        sub = engine.count(_sample(p, L, 0, reg_p != i, reg_s != i), th)
        np.testing.assert_array_equal(sub.npairs[0], full.npairs[0] - full.npairs[1 + i])
        d = full.splus[0] - full.splus[1 + i]
        assert np.abs(sub.splus[0] - d).max() <= 1e-9 * np.abs(full.splus[0]).max(), i


def test_clustered1M_los_permutation(engine, clustered):
    cat, p = clustered
    L = cat.boxsize
    th, _ = _w_thresholds(L)
    a = engine.count(_sample(p, L, 3), th)
    # columns (x, y, z) -> (z, x, y) with LOS 2 -> 0: the projected axes stay (x, y)
    d = dict(cat.data)
    d["Position"] = np.ascontiguousarray(cat.data["Position"][:, [2, 0, 1]])
    d["Position_shape_sample"] = np.ascontiguousarray(cat.data["Position_shape_sample"][:, [2, 0, 1]])
    d["LOS"] = 0
    q = hostprep.prepare(d)
    b = engine.count(_sample(q, L, 3), th)
    np.testing.assert_array_equal(a.npairs[0], b.npairs[0])
    assert np.abs(a.splus[0] - b.splus[0]).max() <= 1e-12 * np.abs(a.splus[0]).max()
    # region (ix, iy, iz) of the original is region (iz, ix, iy) of the permuted box
    n = 3
    perm = [iz * n * n + ix * n + iy for ix in range(n) for iy in range(n) for iz in range(n)]
    np.testing.assert_array_equal(a.npairs[1:], b.npairs[1:][perm])


@pytest.mark.timeout(1200)
def test_clustered1M_against_tree_oracle(engine, clustered):
    """The benchmark count whole against the oracle's tree counter: every
    shape chunk of measure_w_box_jk.py:543-646, summed as :761-780 does."""
    from oracle import post, prep, tree
    cat, p = clustered
    L = cat.boxsize
    th, r_bins = _w_thresholds(L)
    pi_bins = np.linspace(-60, 60, 21)
    res = engine.count(_sample(p, L, 3), th)
    pr = prep.prepare(cat.data)
    jp, js = prep.jackknife_region_indices_by_axis(pr["positions"], pr["shapes"], L, 3)
    ref = tree.count_pairs(pr, r_bins, pi_bins, 0.1, 20.0, L, jk_pos=jp, jk_shape=js, num_box=27,
                           num_nodes=min(16, tree.default_workers()), abs_sums=True, progress=True)
    print("pairs", ref["npairs"], "oracle s", round(ref["seconds"], 1), flush=True)
    assert res.npairs[0].sum() == ref["npairs"] > 3_000_000_000
    np.testing.assert_array_equal(res.dd[0], ref["DD"])
    np.testing.assert_array_equal(res.dd[1:], ref["DD_jk"])
    R = p.R
    sp = res.splus[0] / (2 * R)
    assert _rel(sp, ref["SplusD"], ref["SplusD_abs"]).max() <= 1e-6
    assert _rel(res.scross / (2 * R), ref["ScrossD"], ref["ScrossD_abs"]).max() <= 1e-6
    assert _rel(res.splus[1:], ref["SplusD_jk"], ref["SplusD_jk_abs"]).max() <= 1e-6
    # w_g+ per r_p bin: sum over pi of S+D / RR |dpi| (measure_IA_base.py:553-562),
    # bounded by the same 1e-6 of the summed |terms|
    RR = post.rr_grid("rppi", r_bins, pi_bins, L ** 3, len(pr["positions"]), len(pr["shapes"]))
    w_got = post.project_w(sp / RR, pi_bins)
    w_ref = post.project_w(ref["SplusD"] / RR, pi_bins)
    w_abs = post.project_w(np.maximum(np.abs(ref["SplusD"]), ref["SplusD_abs"]) / RR, pi_bins)
    assert np.all(np.abs(w_got - w_ref) <= 1e-6 * w_abs)
    gg_got = post.project_w(res.dd[0] / RR - 1, pi_bins)
    np.testing.assert_array_equal(gg_got, post.project_w(ref["DD"] / RR - 1, pi_bins))


def _multipole_setup():
    cat = synth.config("multipoles2M")
    r_bins = np.logspace(np.log10(0.1), np.log10(60.0), 16)
    mu_bins = np.linspace(-1, 1, 21)
    th = binning.rmu_thresholds(r_bins, mu_bins, 0.1, 60.0)
    return cat, r_bins, mu_bins, th


def test_multipoles2M_leave_one_region_out(engine):
    cat, r_bins, mu_bins, th = _multipole_setup()
    L = cat.boxsize
    p = hostprep.prepare(cat.data)
    full = engine.count(_sample(p, L, 4), th)
    assert full.npairs[0].sum() > 100_000_000_000
    jk_sub = (L / 2.0) * 2.0 / 4
    reg_p = engine.region_ids(p.positions, 4, jk_sub)
    reg_s = engine.region_ids(p.shapes, 4, jk_sub)
    i = 21
    sub = engine.count(_sample(p, L, 0, reg_p != i, reg_s != i), th)
    np.testing.assert_array_equal(sub.npairs[0], full.npairs[0] - full.npairs[1 + i])
    np.testing.assert_array_equal(sub.dd[0], full.dd[0] - full.dd[1 + i])
    d = full.splus[0] - full.splus[1 + i]
    assert np.abs(sub.splus[0] - d).max() <= 1e-9 * np.abs(full.splus[0]).max()


def test_multipoles2M_shape_chunks_against_tree_oracle(engine):
    """160 seeded shapes (16 chunks of 10) of the multipole box against all
    2,000,000 positions in (r, mu), 64 regions: the oracle's tree counter
    (measure_m_box_jk.py:570-682 per chunk) pair for pair."""
    from oracle import prep, tree
    cat, r_bins, mu_bins, th = _multipole_setup()
    L = cat.boxsize
    pr = prep.prepare(cat.data)
    jp, js = prep.jackknife_region_indices_by_axis(pr["positions"], pr["shapes"], L, 4)
    chunk = 10
    n_chunks = (len(pr["shapes"]) + chunk - 1) // chunk
    pick = sorted(np.random.default_rng(11).choice(n_chunks, 16, replace=False).tolist())
    sel = np.concatenate([np.arange(c * chunk, min((c + 1) * chunk, len(pr["shapes"]))) for c in pick])
    ref = tree.count_pairs(pr, r_bins, mu_bins, 0.1, 60.0, L, mode="rmu", jk_pos=jp, jk_shape=js,
                           num_box=64, chunk_size=chunk, num_nodes=min(16, tree.default_workers()),
                           chunk_subset=pick, abs_sums=True)
    print("pairs", ref["npairs"], "oracle s", round(ref["seconds"], 1), flush=True)
    p = hostprep.prepare(cat.data)
    res = engine.count(_sample(p, L, 4, None, sel), th)
    assert res.npairs[0].sum() == ref["npairs"] > 10_000_000
    np.testing.assert_array_equal(res.dd[0], ref["DD"])
    np.testing.assert_array_equal(res.dd[1:], ref["DD_jk"])
    assert _rel(res.splus[0] / (2 * p.R), ref["SplusD"], ref["SplusD_abs"]).max() <= 1e-6
    assert _rel(res.scross / (2 * p.R), ref["ScrossD"], ref["ScrossD_abs"]).max() <= 1e-6
    assert _rel(res.splus[1:], ref["SplusD_jk"], ref["SplusD_jk_abs"]).max() <= 1e-6


def test_batch12_equals_single_counts(engine):
    cats = synth.uneven_samples(12, parent=3_000_000, L=205.0, lo=10_000, hi=800_000, shape_frac=0.3,
                                seed=3)
    L = 205.0
    th, _ = _w_thresholds(L)
    samples = [_sample(hostprep.prepare(c.data), L, 3, weighted=True) for c in cats]
    batch = engine.count_batch(samples, th)
    for s, b in zip(samples, batch):
        one = engine.count(s, th)
        np.testing.assert_array_equal(b.npairs, one.npairs)
        assert np.abs(b.dd - one.dd).max() <= 1e-12 * np.abs(one.dd).max()
        assert np.abs(b.splus - one.splus).max() <= 1e-12 * np.abs(one.splus).max()


def test_large20M_shape_chunks_against_tree_oracle(engine):
    from oracle import prep, tree
    cat = synth.config("large20M")
    L = cat.boxsize
    r_bins = np.logspace(np.log10(0.1), np.log10(100.0), 16)
    sb = np.linspace(-100.0, 100.0, 41)
    th = binning.rppi_thresholds(r_bins, sb, 0.1, 100.0)
    pr = prep.prepare(cat.data)
    jp, js = prep.jackknife_region_indices_by_axis(pr["positions"], pr["shapes"], L, 5)
    chunk = 10
    n_chunks = (len(pr["shapes"]) + chunk - 1) // chunk
    pick = sorted(np.random.default_rng(5).choice(n_chunks, 16, replace=False).tolist())
    sel = np.concatenate([np.arange(c * chunk, min((c + 1) * chunk, len(pr["shapes"]))) for c in pick])
    ref = tree.count_pairs(pr, r_bins, sb, 0.1, 100.0, L, jk_pos=jp, jk_shape=js, num_box=125,
                           chunk_size=chunk, num_nodes=min(16, tree.default_workers()), chunk_subset=pick,
                           abs_sums=True)
    print("pairs", ref["npairs"], "oracle s", round(ref["seconds"], 1), flush=True)
    p = hostprep.prepare(cat.data)
    res = engine.count(_sample(p, L, 5, None, sel), th)
    assert res.npairs[0].sum() == ref["npairs"] > 100_000_000
    np.testing.assert_array_equal(res.dd[0], ref["DD"])
    np.testing.assert_array_equal(res.dd[1:], ref["DD_jk"])
    assert _rel(res.splus[0] / (2 * p.R), ref["SplusD"], ref["SplusD_abs"]).max() <= 1e-6
    assert _rel(res.scross / (2 * p.R), ref["ScrossD"], ref["ScrossD_abs"]).max() <= 1e-6
    assert _rel(res.splus[1:], ref["SplusD_jk"], ref["SplusD_jk_abs"]).max() <= 1e-6
