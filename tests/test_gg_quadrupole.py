"""The gg quadrupole extension (MeasureIABox.measure_gg_quadrupole), CPU only.

Pinned by closed forms on the mu grid the reference uses (20 and 200 bins):
a grid xi(r, mu) = B(r) P_2(mu) gives xi_2 = B(r) up to the midpoint rule's
O(dmu^2) error, and a constant grid c gives exactly (5/2) c (-dmu^2/4): the
midpoint rule misses the integral of a quadratic by (b-a) dmu^2 f''/24, and
the integral of P_2 over [-1, 1] is 0.
"""
import numpy as np
import pytest

from measure_ia_amd import MeasureIABox
from measure_ia_amd.hdf5io import H5Store


def _box(fn, nmu):
    d = {"Position": np.zeros((1, 3)), "Position_shape_sample": np.zeros((1, 3)),
         "Axis_Direction": np.ones((1, 2)), "LOS": 2, "q": np.ones(1)}
    return MeasureIABox(d, fn, None, 99, [0.1, 60.0], 15, nmu, None, boxsize=205.0, verbose=False)


@pytest.mark.parametrize("nmu, tol", [(20, 1.5e-2), (200, 1.5e-4)])
def test_quadrupole_closed_forms(tmp_path, nmu, tol):
    fn = str(tmp_path / "q.hdf5")
    box = _box(fn, nmu)
    mu_c = box.mu_r_bins[:-1] + np.diff(box.mu_r_bins) / 2
    B = np.linspace(0.5, 2.0, 15)
    p2 = 0.5 * (3 * mu_c ** 2 - 1)
    g = "Snapshot_99/multipoles/xi_gg"
    K = 3
    with H5Store(fn, "a") as st:
        st.write(g, "d", B[:, None] * p2[None, :])
        st.write(g, "c", np.full((15, nmu), 7.0))
        st.write(g, "c_mu_r", mu_c)
        st.write(g, "d_mu_r", mu_c)
        for i in range(K):
            st.write(f"{g}/d_jk{K}", f"d_{i}", (1 + 0.1 * i) * B[:, None] * p2[None, :])
            st.write(f"{g}/d_jk{K}", f"d_{i}_mu_r", mu_c)
    box.measure_gg_quadrupole("d", num_jk=K)
    with H5Store(fn, "r") as st:
        q = st.read("Snapshot_99/multipoles_gg_l2/d")
        reals = [st.read(f"Snapshot_99/multipoles_gg_l2/d_jk{K}/d_{i}") for i in range(K)]
        mean = st.read(f"Snapshot_99/multipoles_gg_l2/d_mean_{K}")
        cov = st.read(f"Snapshot_99/multipoles_gg_l2/d_jackknife_cov_{K}")
    np.testing.assert_allclose(q, B, rtol=tol)
    h = 2.0 / nmu
    qc = box.measure_gg_quadrupole("c", return_output=True)[:, 1]
    np.testing.assert_allclose(qc, 2.5 * 7.0 * (-h * h / 4), rtol=1e-9)
    for i in range(K):
        np.testing.assert_allclose(reals[i], (1 + 0.1 * i) * B, rtol=tol)
    np.testing.assert_allclose(mean, 1.1 * B, rtol=tol)
    # realisations (0.9 + 0.1 i) B around the mean: cov = (n-1)/n sum d d^T
    d = np.array(reals) - mean
    np.testing.assert_allclose(cov, (K - 1) / K * d.T @ d, rtol=1e-12, atol=1e-15)
    out = box.measure_gg_quadrupole("d", return_output=True)
    np.testing.assert_allclose(out[:, 1], q)
