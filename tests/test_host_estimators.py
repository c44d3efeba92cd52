"""Product host post-processing against the reference's fixtures and the oracle."""
import numpy as np
import pytest

from measure_ia_amd import estimators as E
from oracle import post, prep

L = 205.0
N = {"fixture_small": {"All": 766, "high": 292, "low": 474},
     "fixture_large": {"All": 4547, "high": 766, "low": 3781}}


def _bins():
    return prep.bin_edges(0.1, 20.0, 10, 8, None, L)


@pytest.mark.parametrize("fx", ["fixture_small", "fixture_large"])
@pytest.mark.parametrize("ds", ["All", "high", "low"])
def test_rr_against_fixture(request, fx, ds):
    f = request.getfixturevalue(fx)
    n = N[fx][ds]
    r_bins, pi_bins, mu_bins = _bins()
    np.testing.assert_array_equal(E.rr_rppi(r_bins, pi_bins, L ** 3, n, n),
                                  f[f"Snapshot_99/w/xi_gg/{ds}_RR_gg"])
    np.testing.assert_allclose(E.rr_rmu(r_bins, mu_bins, L ** 3, n, n),
                               f[f"Snapshot_99/multipoles/xi_gg/{ds}_RR_gg"], rtol=1e-14)


@pytest.mark.parametrize("fx", ["fixture_small", "fixture_large"])
@pytest.mark.parametrize("ds", ["All", "high", "low"])
def test_projections_and_covariance(request, fx, ds):
    f = request.getfixturevalue(fx)
    r_bins, pi_bins, mu_bins = _bins()
    for corr in ("g_plus", "gg"):
        xi = f[f"Snapshot_99/w/xi_{corr}/{ds}"]
        pi = f[f"Snapshot_99/w/xi_{corr}/{ds}_pi"]
        np.testing.assert_array_equal(E.w_projection(xi, pi_bins, pi), f[f"Snapshot_99/w_{corr}/{ds}"])
        l = 2 if corr == "g_plus" else 0
        xm = f[f"Snapshot_99/multipoles/xi_{corr}/{ds}"]
        mu = f[f"Snapshot_99/multipoles/xi_{corr}/{ds}_mu_r"]
        got = E.multipole(xm, mu, mu_bins, l, l)
        ref = f[f"Snapshot_99/multipoles_{corr}/{ds}"]
        assert np.all(np.abs(got - ref) <= 1e-13 * np.sum(np.abs(xm), axis=1) + 1e-300)
    for grp in ("w_g_plus", "w_gg", "multipoles_g_plus", "multipoles_gg"):
        reals = [f[f"Snapshot_99/{grp}/{ds}_{i}"] for i in range(8)]
        mean, std, cov = E.jackknife_combine(reals, 10)
        np.testing.assert_array_equal(mean, f[f"Snapshot_99/{grp}/{ds}_mean_8"])
        np.testing.assert_array_equal(std, f[f"Snapshot_99/{grp}/{ds}_jackknife_8"])
        np.testing.assert_array_equal(cov, f[f"Snapshot_99/{grp}/{ds}_jackknife_cov_8"])


def test_pi_centre_mismatch_raises():
    r_bins, pi_bins, _ = _bins()
    xi = np.ones((10, 8))
    with pytest.raises(ValueError):
        E.w_projection(xi, pi_bins, np.linspace(-1, 1, 8))


def test_legendre_closed_forms_match_scipy():
    from scipy.special import lpmn
    import warnings
    mu = np.linspace(-0.975, 0.975, 40)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", DeprecationWarning)
        for l, m in ((0, 0), (2, 2), (2, 0)):
            ref = np.array([lpmn(m, l, x)[0][-1, -1] for x in mu])
            np.testing.assert_allclose(E.legendre_assoc(l, m, mu), ref, rtol=1e-14, atol=1e-15)


def test_estimates_match_oracle_estimator():
    """Realisations, RR_jk and R_jk of the product against oracle.post on the same counts."""
    rng = np.random.default_rng(4)
    r_bins, pi_bins, _ = prep.bin_edges(0.1, 20.0, 10, 20, 60.0, L)
    K = 27
    counts = {"DD": rng.integers(0, 100, (10, 20)).astype(float),
              "SplusD": rng.normal(size=(10, 20)), "ScrossD": rng.normal(size=(10, 20)),
              "variance": rng.uniform(size=(10, 20)),
              "DD_jk": rng.integers(0, 10, (K, 10, 20)).astype(float),
              "SplusD_jk": rng.normal(size=(K, 10, 20))}
    n_pos, n_shape = 5000, 2000
    jk_pos = rng.integers(0, K, n_pos)
    jk_shape = rng.integers(0, K, n_shape)
    ws = rng.uniform(0.5, 1.5, n_shape)
    e = rng.uniform(0, 0.5, n_shape)
    R = sum(ws * (1 - e ** 2 / 2.0)) / sum(ws)
    R_jk = prep.responsivity_jk(ws, e, jk_shape, K)
    ref = post.estimator("rppi", counts, r_bins, pi_bins, L, n_pos, n_shape, R,
                         jk=dict(jk_pos=jk_pos, jk_shape=jk_shape, num_box=K, R_jk=R_jk))
    rc = np.stack([np.bincount(jk_pos, minlength=K), np.bincount(jk_shape, minlength=K)], axis=1)
    ra = np.bincount(jk_shape, weights=ws * (1 - e ** 2 / 2.0), minlength=K)
    rw = np.bincount(jk_shape, weights=ws, minlength=K)
    Rj = E.responsivity_jk(sum(ws * (1 - e ** 2 / 2.0)), sum(ws), ra, rw)
    np.testing.assert_allclose(Rj, R_jk, rtol=1e-13)
    est = E.Estimates("rppi", r_bins, pi_bins, L, n_pos, n_shape, R, counts, rc, Rj, K,
                      variance=counts["variance"])
    np.testing.assert_array_equal(est.RR, ref["RR"])
    np.testing.assert_array_equal(est.xi_gg, ref["xi_gg"])
    np.testing.assert_array_equal(est.xi_g_plus, ref["xi_g_plus"])
    np.testing.assert_array_equal(est.sigmasq, ref["sigmasq"])
    np.testing.assert_array_equal(est.RR_jk, ref["RR_jk"])
    for i in range(K):
        np.testing.assert_allclose(est.real_xi_g_plus[i], ref[("jk", i)]["xi_g_plus"], rtol=1e-13)
        np.testing.assert_allclose(est.real_SplusD[i], ref[("jk", i)]["SplusD"], rtol=1e-13)
        np.testing.assert_array_equal(est.real_xi_gg[i], ref[("jk", i)]["xi_gg"])
        np.testing.assert_array_equal(est.real_DD[i], ref[("jk", i)]["DD"])


def test_box_static_helpers():
    """measure_IA_base.py:186-227, 353-401: the helpers a MeasureIABox exposes."""
    from measure_ia_amd import MeasureIABox
    rng = np.random.default_rng(4)
    a, b = rng.normal(size=(50, 3)), rng.normal(size=(50, 3))
    dot = MeasureIABox.calculate_dot_product_arrays(a, b)
    np.testing.assert_array_equal(dot, (a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1]) + a[:, 2] * b[:, 2])
    ep, ex = MeasureIABox.get_ellipticity(np.array([0.5, 0.5, 0.5]), np.array([0.0, np.pi / 4, np.pi / 2]))
    np.testing.assert_allclose(ep, [0.5, 0.0, -0.5], atol=1e-16)
    np.testing.assert_allclose(ex, [0.0, 0.5, 0.0], atol=1e-16)
    d = MeasureIABox.setdiff2D([[1, 2, 3], [4, 5]], [[2], [5, 6]])
    assert [list(x) for x in d] == [[1, 3], [4]]
    d = MeasureIABox.setdiff_omit([[1, 2, 3], [4, 5], [6, 7]], [2, 6], [0, 2])
    assert [list(x) for x in d] == [[1, 3], [7]]
    with pytest.raises(AssertionError):
        MeasureIABox.setdiff2D([[1]], [[1], [2]])


def test_read_output_keeps_values_when_covariance_missing(tmp_path):
    """read_data.py:336-374: assignments inside each try happen one by one."""
    from measure_ia_amd import ReadData
    from measure_ia_amd.hdf5io import H5Store
    fn = tmp_path / "cat.hdf5"
    with H5Store(str(fn), "a") as st:
        st.write("Snapshot_99/w_gg", "All", np.arange(4.0))
        st.write("Snapshot_99/w_gg", "All_rp", np.arange(4.0) + 1)
    rd = ReadData("TNG300", "cat", 99, data_path=str(tmp_path))
    rd.read_MeasureIA_output("All", 27)
    np.testing.assert_array_equal(rd.w_gg, np.arange(4.0))
    np.testing.assert_array_equal(rd.rp, np.arange(4.0) + 1)
    assert rd.cov_w_gg is None and rd.w_gp is None and rd.multipoles_gg is None


@pytest.mark.parametrize("est", ["clusters", "galaxies"])
def test_lightcone_estimator_table_matches_restatement(est):
    """estimators.lightcone_xi (the product's table) against the test helper's
    independent restatement of measure_IA_base.py:670-743, bit for bit."""
    from measure_ia_amd import estimators
    from reflc import _estimate
    rng = np.random.default_rng(5)
    cnt = {k: rng.uniform(1.0, 50.0, (4, 6)) for k in ("SpD", "DD", "SpR", "SR", "RD", "RR")}
    ns = {"S": 1234, "D": 5678, "R_S": 23456, "R_D": 34567}
    names = {"S+D": "SpD", "DD": "DD", "S+R": "SpR", "SR": "SR", "RD": "RD", "RR": "RR"}
    gp = estimators.lightcone_xi(est, "g+", lambda n: cnt[names[n]].copy(), ns)
    gg = estimators.lightcone_xi(est, "gg", lambda n: cnt[names[n]].copy(), ns)
    xgp, xgg = _estimate(est, cnt, ns)
    np.testing.assert_array_equal(gp, xgp)
    np.testing.assert_array_equal(gg, xgg)
    # a planted slip (galaxies' +1 where clusters have -1) fails
    other = "galaxies" if est == "clusters" else "clusters"
    assert not np.allclose(estimators.lightcone_xi(other, "gg", lambda n: cnt[names[n]].copy(), ns), xgg)
