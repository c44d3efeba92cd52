"""Cross-dataset jackknife covariance (measure_jackknife.py:485-648).

Known answers: the reference's own (disabled) tests
tests/test_covariance_multiple_datasets.py:10-40 and
tests/test_full_cov_matrix_projections.py:5-14 state the covariances of three
4-bin datasets with 3 realisations; their fixture file is not shipped, so the
realisations below are built to the same deviations (set2 = set1 with
realisations 0 and 1 exchanged, set3 = 2 x set2), which reproduces every
matrix those tests state.
"""
import numpy as np
import pytest

from measure_ia_amd import MeasureIABox
from measure_ia_amd.hdf5io import H5Store

D1 = np.array([[1, 1, 1, -1], [0, -1, -1, 1], [-1, 0, 0, 0]], float)
MEANS = {"set1": np.array([0.5, -0.25, 2.0, 1.0]), "set2": np.array([1.0, 0.0, -1.0, 0.5]),
         "set3": np.array([0.0, 3.0, 0.25, -2.0])}
DEV = {"set1": D1, "set2": D1[[1, 0, 2]], "set3": 2 * D1[[1, 0, 2]]}


def _box(tmp_path):
    fn = str(tmp_path / "cov.hdf5")
    with H5Store(fn, "a") as st:
        for ds in ("set1", "set2", "set3"):
            for b in range(3):
                st.write(f"Snapshot_99/w_g_plus/{ds}_jk3", f"{ds}_{b}", MEANS[ds] + DEV[ds][b])
    data = {"Position": np.zeros((2, 3)), "Position_shape_sample": np.zeros((2, 3)),
            "Axis_Direction": np.ones((2, 2)), "LOS": 2, "q": np.ones(2)}
    return MeasureIABox(data, fn, None, 99, [0.1, 20.0], 4, 4, None, boxsize=100.0, verbose=False)


def test_covariance_multiple_datasets_known_answers(tmp_path):
    box = _box(tmp_path)
    c = lambda names: box.measure_covariance_multiple_datasets("w_g_plus", names, 3, True)[0]
    cov1, cov2, cov3 = c(["set1"]), c(["set2"]), c(["set3"])
    np.testing.assert_array_equal(cov1, cov2)
    np.testing.assert_array_equal(cov2 * 4, cov3)
    np.testing.assert_allclose(
        cov1, 2 / 3 * np.array([[2, 1, 1, -1], [1, 2, 2, -2], [1, 2, 2, -2], [-1, -2, -2, 2]]),
        rtol=1e-15, atol=1e-15)
    cov12 = c(["set1", "set2"])
    np.testing.assert_allclose(
        cov12, 2 / 3 * np.array([[1, -1, -1, 1], [-1, -2, -2, 2], [-1, -2, -2, 2], [1, 2, 2, -2]]),
        rtol=1e-15, atol=1e-15)
    np.testing.assert_array_equal(c(["set1", "set3"]), 2 * cov12)
    np.testing.assert_array_equal(c(["set2", "set3"]), 2 * cov1)
    # std: sqrt of (n-1)/n times the running sum; the cross "std" is NaN where negative
    _, std1 = box.measure_covariance_multiple_datasets("w_g_plus", ["set1"], 3, True)
    np.testing.assert_allclose(std1, np.sqrt(np.diag(cov1)), rtol=1e-15)
    _, std12 = box.measure_covariance_multiple_datasets("w_g_plus", ["set1", "set2"], 3, True)
    assert np.isnan(std12[1]) and std12[0] == pytest.approx(np.sqrt(2 / 3))
    with pytest.raises(KeyError):
        box.measure_covariance_multiple_datasets("w_g_plus", ["set1", "set2", "set3"], 3, True)
    with pytest.raises(ValueError):
        box.measure_covariance_multiple_datasets("xi_gg", ["set1"], 3, True)


def test_full_cov_matrix_projections(tmp_path):
    box = _box(tmp_path)
    for ds in ("set1", "set2", "set3"):
        box.measure_covariance_multiple_datasets("w_g_plus", [ds], 3, False)
    cov3, c12, c13, c23 = box.create_full_cov_matrix_projections("w_g_plus", ["set1", "set2", "set3"], 3,
                                                                 True)
    # the relations of tests/test_full_cov_matrix_projections.py:9-14
    np.testing.assert_array_equal(cov3[0:8, 0:8], c12)
    np.testing.assert_array_equal(cov3[4:12, 4:12], c23)
    np.testing.assert_array_equal(cov3[0:4, 0:4], c13[0:4, 0:4])
    np.testing.assert_array_equal(cov3[8:12, 0:4], c13[4:8, 0:4])
    np.testing.assert_array_equal(cov3[8:12, 8:12], c13[4:8, 4:8])
    np.testing.assert_array_equal(cov3[0:4, 8:12], c13[0:4, 4:8])
    # the reference's block placement: (set2, set3) cross covariance in the x-z slot
    cov_23 = box.measure_covariance_multiple_datasets("w_g_plus", ["set2", "set3"], 3, True)[0]
    cov_13 = box.measure_covariance_multiple_datasets("w_g_plus", ["set1", "set3"], 3, True)[0]
    np.testing.assert_array_equal(cov3[0:4, 8:12], cov_23)
    np.testing.assert_array_equal(cov3[4:8, 8:12], cov_13)
    # written datasets
    box.create_full_cov_matrix_projections("w_g_plus", ["set1", "set2", "set3"], 3, False)
    with H5Store(box.output_file_name, "r") as st:
        np.testing.assert_array_equal(
            st.read("Snapshot_99/w_g_plus/set1_set2_set3_combined_jackknife_cov_3"), cov3)
        np.testing.assert_array_equal(st.read("Snapshot_99/w_g_plus/set1_set2_jackknife_cov_3"), c12[0:4, 4:8])


def test_lightcone_class_has_the_covariance_methods(tmp_path):
    """The reference's MeasureIALightcone inherits both methods from
    MeasureJackknife; lightcone outputs carry no snapshot group."""
    from measure_ia_amd import MeasureIALightcone
    fn = str(tmp_path / "lc_cov.hdf5")
    with H5Store(fn, "a") as st:
        for ds in ("set1", "set2", "set3"):
            for b in range(3):
                st.write(f"w_g_plus/{ds}_jk3", f"{ds}_{b}", MEANS[ds] + DEV[ds][b])
    cols = {k: np.zeros(2) for k in ("RA", "DEC", "Redshift", "RA_shape_sample", "DEC_shape_sample",
                                     "Redshift_shape_sample", "e1", "e2")}
    lc = MeasureIALightcone(cols, {k: np.zeros(2) for k in ("RA", "DEC", "Redshift")}, fn,
                            separation_limits=[0.1, 20.0], num_bins_r=4, num_bins_pi=4, pi_max=10.0)
    cov1 = lc.measure_covariance_multiple_datasets("w_g_plus", ["set1"], 3, True)[0]
    np.testing.assert_allclose(
        cov1, 2 / 3 * np.array([[2, 1, 1, -1], [1, 2, 2, -2], [1, 2, 2, -2], [-1, -2, -2, 2]]),
        rtol=1e-15, atol=1e-15)
    for ds in ("set1", "set2", "set3"):
        lc.measure_covariance_multiple_datasets("w_g_plus", [ds], 3, False)
    cov3, c12, _, c23 = lc.create_full_cov_matrix_projections("w_g_plus", ["set1", "set2", "set3"], 3, True)
    np.testing.assert_array_equal(cov3[0:8, 0:8], c12)
    np.testing.assert_array_equal(cov3[4:12, 4:12], c23)
