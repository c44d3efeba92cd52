"""Jackknife covariance across datasets, from realisations in an output file.

Restates MeasureJackknife.measure_covariance_multiple_datasets and
create_full_cov_matrix_projections (src/measureia/measure_jackknife.py:485-648).
The reference defines them on its lightcone class, but they only read the
jackknife realisations ``{snap_group}/{corr_type}/{ds}_jk{n}/{ds}_{b}`` that the
box measurements write too, so MeasureIABox exposes them for box outputs.

Arithmetic follows the reference's order (running sums over realisations, then
the (n-1)/n factor of Singh 2023), so results match it bit for bit.  Two
quirks are kept: the cross "std" of two datasets is sqrt((n-1)/n * sum dx*dy),
NaN where that sum is negative; and create_full_cov_matrix_projections loads
the (d0, d2) cross covariance into the block it calls y-z and the (d1, d2)
one into the block it calls x-z (measure_jackknife.py:600-602), so the
combined matrices carry those blocks where the reference puts them.
"""
from __future__ import annotations

import numpy as np

from .hdf5io import H5Store

VALID_CORR = ("w_g_plus", "multipoles_g_plus", "w_gg", "multipoles_gg")


def _jk_mean(st, base, ds, num_box, nr):
    mean = np.zeros(nr)
    for b in range(num_box):
        mean += st.read(f"{base}/{ds}_jk{num_box}/{ds}_{b}")
    mean /= num_box
    return mean


def measure_covariance_multiple_datasets(filename, snap_group, corr_type, dataset_names, num_box,
                                         num_bins_r, return_output=False):
    """Covariance and std of one dataset with itself or of two datasets (:485-577)."""
    if corr_type not in VALID_CORR:
        raise ValueError("corr_type must be 'w_g_plus', 'w_gg', 'multipoles_g_plus' or 'multipoles_gg'.")
    nr = num_bins_r
    base = f"{snap_group}/{corr_type}"
    with H5Store(filename, "a") as st:
        means = [_jk_mean(st, base, ds, num_box, nr) for ds in dataset_names]
        cov = np.zeros((nr, nr))
        std = np.zeros(nr)
        if len(dataset_names) == 1:
            ds = dataset_names[0]
            for b in range(num_box):
                x = st.read(f"{base}/{ds}_jk{num_box}/{ds}_{b}")
                dx = x - means[0]
                std += dx ** 2
                for i in range(nr):
                    cov[:, i] += dx * (x[i] - means[0][i])
        elif len(dataset_names) == 2:
            d0, d1 = dataset_names
            for b in range(num_box):
                x = st.read(f"{base}/{d0}_jk{num_box}/{d0}_{b}")
                y = st.read(f"{base}/{d1}_jk{num_box}/{d1}_{b}")
                std += (x - means[0]) * (y - means[1])
                for i in range(nr):
                    cov[:, i] += (x - means[0]) * (y[i] - means[1][i])
        else:
            raise KeyError("Too many datasets given, choose either 1 or 2")
        std *= (num_box - 1) / num_box
        with np.errstate(invalid="ignore"):
            std = np.sqrt(std)
        cov *= (num_box - 1) / num_box
        if filename is not None and not return_output:
            name = "_".join(dataset_names)
            st.write(base, f"{name}_jackknife_cov_{num_box}", cov)
            st.write(base, f"{name}_jackknife_{num_box}", std)
            return None
    return cov, std


def create_full_cov_matrix_projections(filename, snap_group, corr_type, dataset_names, num_box,
                                       num_bins_r, return_output=False):
    """Covariance of 3 projections and of every pair of them (:579-648)."""
    d0, d1, d2 = dataset_names[:3]
    for pair in ((d0, d1), (d0, d2), (d1, d2)):
        measure_covariance_multiple_datasets(filename, snap_group, corr_type, list(pair), num_box,
                                             num_bins_r)
    base = f"{snap_group}/{corr_type}"
    with H5Store(filename, "a") as st:
        cov_xx = st.read(f"{base}/{d0}_jackknife_cov_{num_box}")
        cov_yy = st.read(f"{base}/{d1}_jackknife_cov_{num_box}")
        cov_zz = st.read(f"{base}/{d2}_jackknife_cov_{num_box}")
        cov_xy = st.read(f"{base}/{d0}_{d1}_jackknife_cov_{num_box}")
        cov_yz = st.read(f"{base}/{d0}_{d2}_jackknife_cov_{num_box}")
        cov_xz = st.read(f"{base}/{d1}_{d2}_jackknife_cov_{num_box}")
        cov3 = np.block([[cov_xx, cov_xy, cov_xz],
                         [cov_xy.T, cov_yy, cov_yz],
                         [cov_xz.T, cov_yz.T, cov_zz]])
        cov2xy = np.block([[cov_xx, cov_xy], [cov_xy.T, cov_yy]])
        cov2xz = np.block([[cov_xx, cov_xz], [cov_xz.T, cov_zz]])
        cov2yz = np.block([[cov_yy, cov_yz], [cov_yz.T, cov_zz]])
        if return_output:
            return cov3, cov2xy, cov2xz, cov2yz
        st.write(base, f"{d0}_{d1}_{d2}_combined_jackknife_cov_{num_box}", cov3)
        st.write(base, f"{d0}_{d1}_combined_jackknife_cov_{num_box}", cov2xy)
        st.write(base, f"{d0}_{d2}_combined_jackknife_cov_{num_box}", cov2xz)
        st.write(base, f"{d1}_{d2}_combined_jackknife_cov_{num_box}", cov2yz)
    return None
