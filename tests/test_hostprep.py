"""Input preparation, jackknife regions and SimInfo (CPU)."""
import numpy as np
import pytest

from measure_ia_amd import hostprep
from measure_ia_amd.box import MeasureIABox, _cube_root
from measure_ia_amd.siminfo import SimInfo
from oracle import prep


def _data(n=500, seed=0, weights=False):
    rng = np.random.default_rng(seed)
    d = {"Position": rng.uniform(0, 50, (n, 3)), "Position_shape_sample": rng.uniform(0, 50, (n // 2, 3)),
         "Axis_Direction": rng.normal(size=(n // 2, 2)), "LOS": 2, "q": rng.uniform(0.2, 1, n // 2)}
    if weights:
        d["weight"] = rng.uniform(0.5, 1.5, n)
        d["weight_shape_sample"] = rng.uniform(0.5, 1.5, n // 2)
    return d


@pytest.mark.parametrize("ell", ["distortion", "ellipticity"])
@pytest.mark.parametrize("weights", [False, True])
def test_prepare_matches_oracle(ell, weights):
    d = _data(weights=weights)
    a = hostprep.prepare(d, None, ell)
    b = prep.prepare(d, None, ell)
    assert a.R == b["R"]  # bit-exact, sequential sum like Python's builtin
    np.testing.assert_array_equal(a.e, b["e"])
    np.testing.assert_array_equal(a.axis, b["axis"])
    np.testing.assert_array_equal(a.not_los, b["not_los"])


def test_prepare_masks_quirk():
    """A missing weight mask selects the FIRST k weights (measure_w_box_jk.py:113-122)."""
    d = _data(weights=True)
    rng = np.random.default_rng(3)
    m = {"Position": rng.uniform(size=500) < 0.5, "Position_shape_sample": rng.uniform(size=250) < 0.5}
    m["Axis_Direction"] = m["Position_shape_sample"]
    m["q"] = m["Position_shape_sample"]
    m2 = dict(m)
    a = hostprep.prepare(d, m, "distortion")
    b = prep.prepare(d, m2, "distortion")
    k = int(m["Position"].sum())
    np.testing.assert_array_equal(a.weight, d["weight"][:k])
    np.testing.assert_array_equal(a.weight, b["weight"])
    assert "weight" in m  # stored back into the masks dict, as the reference does
    assert a.R == b["R"]


def test_bad_ellipticity_raises():
    with pytest.raises(ValueError):
        hostprep.prepare(_data(), None, "bogus")


def test_jackknife_known_answer():
    """tests/test_w_jk.py:16-23 of the reference: 2^3 split of a 3-box."""
    com = np.array([[1, 1, 1], [2, 1, 2], [2.5, 2.5, 1.51], [1, 2, 2]], dtype=float)
    got = hostprep.jackknife_regions(com, 3.0, 2)
    np.testing.assert_array_equal(got, [0, 5, 7, 3])


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_jackknife_vectorised_equals_loop(n):
    rng = np.random.default_rng(n)
    L = 9.0
    x = rng.uniform(0, L, (4000, 3))
    x[:400] = np.round(x[:400] * 2) / 2  # faces of many splits
    x[400:420] = 0.0
    x[420:440] = L
    x[440:460, 1] = L / n * np.arange(20) % L  # exact face multiples
    ref, _ = prep.jackknife_region_indices(x, x[:1], L, n)
    np.testing.assert_array_equal(hostprep.jackknife_regions(x, L, n), ref)


def test_cube_root():
    assert [_cube_root(k) for k in (1, 8, 27, 64, 125)] == [1, 2, 3, 4, 5]
    with pytest.raises(ValueError):
        _cube_root(10)


def test_sim_input():
    """tests/test_sim_input.py:4-46 of the reference."""
    v = SimInfo("TNG300", 30)
    assert (v.simname, v.snapshot, v.boxsize, v.snap_group) == ("TNG300", "30", 205.0, "Snapshot_30/")
    v = SimInfo("TNG300", None)
    assert (v.snapshot, v.boxsize, v.snap_group) == (None, 205.0, "")
    v = SimInfo(None, 30, boxsize=300.0)
    assert (v.simname, v.snapshot, v.boxsize, v.snap_group) == (None, "30", 300.0, "Snapshot_30/")
    v = SimInfo(None, None)
    assert v.boxsize is None and v.snap_group == ""
    v = SimInfo("TNG300", 30, boxsize=300.0)
    assert v.boxsize == 205.0
    assert SimInfo("FLAMINGO_L1", None).boxsize == 1000.0 * 0.681
    with pytest.raises(KeyError):
        SimInfo("FLAMINGO", None)


def test_front_end_construction_and_validation():
    d = _data()
    box = MeasureIABox(d, None, None, None, [0.1, 20.0], 10, 8, None, boxsize=50.0, verbose=False)
    assert box.pi_max == 25.0
    np.testing.assert_array_equal(box.pi_bins, np.linspace(-25.0, 25.0, 9))
    np.testing.assert_array_equal(box.r_bins, np.logspace(np.log10(0.1), np.log10(20.0), 11))
    assert "weight" in d and len(d["weight"]) == 500
    with pytest.raises(ValueError):
        MeasureIABox(d, None, None, None, boxsize=None, verbose=False)
    with pytest.raises(ValueError):
        MeasureIABox(d, None, boxsize=50.0, backend="cpu", verbose=False)
    # bad corr_type and bad num_jk fail before any device work
    with pytest.raises(KeyError):
        box.measure_xi_w("x", "g-", 8, temp_file_path=None)
    with pytest.raises(ValueError):
        box.measure_xi_w("x", "both", 9, temp_file_path=None)
    mixed = MeasureIABox(None, None, "TNG300", 99, verbose=False)
    assert mixed.Num_position == 0


def test_lightcone_class_has_the_base_helpers():
    """MeasureIALightcone carries MeasureIABase's helpers, as the reference's
    lightcone class inherits them (measure_IA_base.py:186-401)."""
    from measure_ia_amd import MeasureIABox, MeasureIALightcone
    cols = {k: np.zeros(2) for k in ("RA", "DEC", "Redshift", "RA_shape_sample", "DEC_shape_sample",
                                     "Redshift_shape_sample", "e1", "e2")}
    lc = MeasureIALightcone(cols, {k: np.zeros(2) for k in ("RA", "DEC", "Redshift")}, None,
                            separation_limits=[0.1, 20.0], num_bins_r=4, num_bins_pi=4, pi_max=10.0)
    a = np.arange(6.0).reshape(3, 2)
    np.testing.assert_array_equal(lc.calculate_dot_product_arrays(a, a), [1.0, 13.0, 41.0])
    assert [list(x) for x in lc.setdiff2D([[1, 2, 3], [4, 5]], [[2], [5]])] == [[1, 3], [4]]
    assert [list(x) for x in lc.setdiff_omit([[1, 2], [3, 4]], [2, 3], [1])] == [[4]]
    assert lc.get_random_pairs(2.0, 1.0, 5.0, -5.0, 1000.0, "cross", 10, 4) == \
        MeasureIABox.get_random_pairs(2.0, 1.0, 5.0, -5.0, 1000.0, "cross", 10, 4)
    assert lc.get_volume_spherical_cap(0.5, 2.0) == MeasureIABox.get_volume_spherical_cap(0.5, 2.0)
    v = lc.get_random_pairs_r_mur(2.0, 1.0, 1.0, -1.0, 1000.0, "cross", 10, 4)
    assert v == pytest.approx(10 * 4 * (4 / 3 * np.pi * (8 - 1)) / 1000.0 * 9 / 10, rel=1e-12)
