"""Lightcone counter (mia_count_lightcone) against the FP64 oracle.

Every case: integer pair counts of the total and of every patch tally equal;
weighted DD within 1e-12 relative; S+D and SxD within 1e-6 of
max(|ref|, sum |terms|) per bin (the same bar as the box counter).
"""
import numpy as np
import pytest

from lccases import CASES, bins, device_arrays, make_case, oracle_counts

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    from measure_ia_amd.engine import get_engine
    return get_engine(0)


def _rel(got, ref, scale):
    den = np.maximum(np.abs(ref), scale)
    d = np.abs(got - ref)
    return np.where(den > 0, d / np.where(den > 0, den, 1), d)


@pytest.mark.parametrize("name", CASES)
def test_lightcone_counts(engine, name):
    a, cfg = make_case(name)
    _, _, th = bins(cfg)
    print(f"case {name}", flush=True)
    res = engine.count_lightcone(device_arrays(a, cfg), th)
    ref = oracle_counts(a, cfg)
    assert ref["NP"][0].sum() > 1000, "case has too few pairs to test anything"
    np.testing.assert_array_equal(res.npairs, ref["NP"])
    scale = np.abs(ref["DD"]).max()
    assert np.abs(res.dd - ref["DD"]).max() <= 1e-12 * scale
    if cfg["shapes"]:
        assert _rel(res.splus, ref["SplusD"], ref["SplusD_abs"]).max() <= 1e-6
        assert _rel(res.scross, ref["ScrossD"], ref["ScrossD_abs"]).max() <= 1e-6
    else:
        assert res.splus is None and res.scross is None
    st = res.stats
    assert st["candidates"] >= ref["NP"][0].sum()


def test_lightcone_empty_samples(engine):
    a, cfg = make_case("rppi_shapes_patches")
    _, _, th = bins(cfg)
    lc = device_arrays(a, cfg)
    for k in ("ra_s", "dec_s", "chi_s", "weight_shape", "c2", "s2", "patch_s"):
        setattr(lc, k, np.zeros(0))
    res = engine.count_lightcone(lc, th)
    assert res.npairs.sum() == 0 and not res.dd.any() and not res.splus.any()


def test_lightcone_tuning_does_not_change_counts(engine):
    a, cfg = make_case("rppi_shapes_patches")
    _, _, th = bins(cfg)
    lc = device_arrays(a, cfg)
    base = engine.count_lightcone(lc, th)
    for tu in ({"slabs_per_side": 1, "columns_per_side": 1}, {"slabs_per_side": 37, "columns_per_side": 53},
               {"block_shapes": 7, "item_positions": 33, "slabs_per_block": 1}, {"waves": 3},
               {"waves": 1}):
        r = engine.count_lightcone(lc, th, tuning=tu)
        np.testing.assert_array_equal(r.npairs, base.npairs)
        assert np.abs(r.splus - base.splus).max() <= 1e-12 * np.abs(base.splus).max()


def test_lightcone_fullsize_leave_one_patch_out(engine):
    """4,000,000 randoms in a 60 x 30 degree wedge (the bench's RR count, about
    1.6e9 pairs), where the oracle would take days: a fresh count without the
    objects of patch i equals total - tally[i] (measure_jackknife.py:84-130
    builds realisation i that way), exactly for the integer pair counts."""
    from measure_ia_amd import synth
    from measure_ia_amd.engine import LightconeArrays
    rng = np.random.default_rng(13)
    ra, dec, chi = synth.lightcone_points(rng, 4_000_000, ra=(0.0, 60.0), dec=(-15.0, 15.0),
                                          chi=(300.0, 2000.0), clustered_frac=0.0)
    patch = (np.clip((ra / 60.0 * 9).astype(int), 0, 8) * 3 + np.clip(((dec + 15.0) / 30.0 * 3).astype(int), 0, 2))
    patch = patch.astype(np.int32)
    _, _, th = bins(dict(mode="rppi", r_min=0.1, r_max=20.0, nr=10, n2=20, pi_max=60.0))
    cosd = np.cos(dec / 180 * np.pi)
    full = engine.count_lightcone(LightconeArrays(ra=ra, dec=dec, chi=chi, cos_dec=cosd, ra_s=ra, dec_s=dec,
                                                  chi_s=chi, patch=patch, patch_s=patch, n_patches=27), th)
    assert full.npairs[0].sum() > 1_000_000_000
    for i in (0, 13, 26):
        keep = patch != i
        sub = engine.count_lightcone(LightconeArrays(ra=ra[keep], dec=dec[keep], chi=chi[keep], cos_dec=cosd[keep],
                                                     ra_s=ra[keep], dec_s=dec[keep], chi_s=chi[keep]), th)
        np.testing.assert_array_equal(sub.npairs[0], full.npairs[0] - full.npairs[1 + i])
        np.testing.assert_array_equal(sub.dd[0], full.dd[0] - full.dd[1 + i])


def test_fullsize_multipoles_leave_one_patch_out(engine):
    """(r, mu_r) at scale: 2,000,000 randoms in the 60 x 30 degree wedge, 10 log
    r bins over 0.5-40, 20 mu bins, with the multipole counters' r_p cut
    (measure_m_lightcone.py:136-141): a fresh count without patch i equals
    total - tally[i] exactly."""
    from measure_ia_amd import synth
    from measure_ia_amd.engine import LightconeArrays
    from measure_ia_amd.lightcone import rp_cut_threshold
    rng = np.random.default_rng(17)
    ra, dec, chi = synth.lightcone_points(rng, 2_000_000, ra=(0.0, 60.0), dec=(-15.0, 15.0),
                                          chi=(300.0, 2000.0), clustered_frac=0.0)
    patch = (np.clip((ra / 60.0 * 9).astype(int), 0, 8) * 3 + np.clip(((dec + 15.0) / 30.0 * 3).astype(int), 0, 2))
    patch = patch.astype(np.int32)
    _, _, th = bins(dict(mode="rmu", r_min=0.5, r_max=40.0, nr=10, n2=20, pi_max=None))
    cosd = np.cos(dec / 180 * np.pi)
    cut = rp_cut_threshold(1.0)
    full = engine.count_lightcone(LightconeArrays(ra=ra, dec=dec, chi=chi, cos_dec=cosd, ra_s=ra, dec_s=dec,
                                                  chi_s=chi, patch=patch, patch_s=patch, n_patches=27,
                                                  rp2_cut=cut), th)
    assert full.npairs[0].sum() > 100_000_000
    for i in (4, 22):
        keep = patch != i
        sub = engine.count_lightcone(LightconeArrays(ra=ra[keep], dec=dec[keep], chi=chi[keep], cos_dec=cosd[keep],
                                                     ra_s=ra[keep], dec_s=dec[keep], chi_s=chi[keep], rp2_cut=cut),
                                     th)
        np.testing.assert_array_equal(sub.npairs[0], full.npairs[0] - full.npairs[1 + i])


@pytest.mark.parametrize("mode, nr, n2", [("rppi", 60, 60), ("rmu", 100, 100)])
def test_lightcone_large_binning_in_passes(engine, mode, nr, n2):
    """3,600 and 10,000 (r, x) bins with shapes and patches: one wave's
    histogram of the whole binning no longer leaves room for enough waves, so
    the lightcone kernel counts in passes over r-bin ranges.  Counts equal the
    oracle, and explicit pass sizes of 1, 7 and all r bins give the same."""
    a, cfg = make_case("rppi_shapes_patches" if mode == "rppi" else "rmu_shapes_patches")
    cfg.update(nr=nr, n2=n2)
    _, _, th = bins(cfg)
    print(f"large lightcone binning {mode} {nr}x{n2}", flush=True)
    lc = device_arrays(a, cfg)
    res = engine.count_lightcone(lc, th)
    ref = oracle_counts(a, cfg)
    assert ref["NP"][0].sum() > 1000
    np.testing.assert_array_equal(res.npairs, ref["NP"])
    assert np.abs(res.dd - ref["DD"]).max() <= 1e-12 * np.abs(ref["DD"]).max()
    assert _rel(res.splus, ref["SplusD"], ref["SplusD_abs"]).max() <= 1e-6
    assert _rel(res.scross, ref["ScrossD"], ref["ScrossD_abs"]).max() <= 1e-6
    for per_pass in (1, 7, nr):
        r = engine.count_lightcone(lc, th, tuning={"r_bins_per_pass": per_pass})
        np.testing.assert_array_equal(r.npairs, res.npairs)
        assert np.abs(r.splus - res.splus).max() <= 1e-12 * np.abs(res.splus).max()
