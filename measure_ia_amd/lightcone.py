"""Lightcone measurements (MeasureIALightcone) on the HIP lightcone counter.

The reference's lightcone path (src/measureia/measure_IA.py:265-1060,
measure_w_lightcone.py, measure_m_lightcone.py, measure_jackknife.py:59-483)
counts pairs with a brute-force loop over positions, once per data/random
combination and once more per jackknife patch.  Here every counter call is one
device count (``Engine.count_lightcone``) whose per-patch tallies give every
leave-one-patch-out realisation; the arguments, the host-side selection and
the HDF5 datasets are the reference's.

Comoving distances: the reference calls pyccl.  When pyccl is importable a
pyccl Cosmology is used the same way; otherwise ``FlatLCDM`` (flat LambdaCDM
with photons and massless neutrinos, the reference's default parameters)
provides them, and any object with ``comoving_radial_distance(a)`` and
``["h"]`` may be passed as ``cosmology``.
"""
from __future__ import annotations

import math

import numpy as np

from . import binning, estimators
from .box import MeasureIABox
from .engine import LightconeArrays
from .siminfo import SimInfo

_C_KMS = 299792.458


class FlatLCDM:
    """Flat LambdaCDM comoving distances (Mpc) without pyccl.

    Defaults are the reference's fallback cosmology
    ccl.Cosmology(Omega_c=0.225, Omega_b=0.045, sigma8=0.8, h=0.7, n_s=1.0)
    (measure_w_lightcone.py:127-131) with CCL's default radiation: CMB
    photons at T_CMB = 2.725 K and Neff = 3.044 massless neutrinos;
    Omega_Lambda closes the budget.  chi(a) = c/H0 int_a^1 da'/(a'^2 E(a'))
    by the trapezoid rule on a fine grid in ln(a) (2e5 steps), interpolated.  Agreement
    with pyccl itself is not pinned here (pyccl is not installed).
    """

    def __init__(self, Omega_c=0.225, Omega_b=0.045, h=0.7, sigma8=0.8, n_s=1.0, T_CMB=2.725,
                 Neff=3.044):
        self.params = dict(Omega_c=Omega_c, Omega_b=Omega_b, h=h, sigma8=sigma8, n_s=n_s,
                           T_CMB=T_CMB, Neff=Neff)
        self.h = h
        # photon density: Omega_g h^2 = 4 sigma_SB T^4 / c^3 / rho_crit,100
        sigma_sb = 5.670374419e-8
        c = 299792458.0
        G = 6.67430e-11
        H100 = 100e3 / 3.0856775814913673e22
        rho_crit100 = 3 * H100 ** 2 / (8 * math.pi * G)
        og = 4 * sigma_sb * T_CMB ** 4 / c ** 3 / rho_crit100 / h ** 2
        onu = og * Neff * 7.0 / 8.0 * (4.0 / 11.0) ** (4.0 / 3.0)
        self.Om = Omega_c + Omega_b
        self.Or = og + onu
        self.OL = 1.0 - self.Om - self.Or
        self._grid = None

    def __getitem__(self, key):
        if key == "h":
            return self.h
        return self.params[key]

    def _E(self, a):
        return np.sqrt(self.Om / a ** 3 + self.Or / a ** 4 + self.OL)

    def _table(self):
        if self._grid is None:
            x = np.linspace(np.log(1e-4), 0.0, 200001)  # ln a
            a = np.exp(x)
            f = 1.0 / (a * self._E(a))                  # da/(a^2 E) = dx/(a E)
            dx = x[1] - x[0]
            # cumulative integral from ln a to 0 (trapezoid on the fine grid)
            cum = np.concatenate([[0.0], np.cumsum((f[1:] + f[:-1]) * dx / 2)])
            chi = (cum[-1] - cum) * _C_KMS / (100.0 * self.h)
            self._grid = (x, chi)
        return self._grid

    def comoving_radial_distance(self, a):
        """Linear interpolation of the table (uniform in ln a; clamped at its ends)."""
        a = np.asarray(a, dtype=np.float64)
        x, chi = self._table()
        n = len(x)
        t = np.clip((np.log(a) - x[0]) / (x[1] - x[0]), 0.0, n - 1.0)
        i = np.minimum(t.astype(np.int64), n - 2)
        f = t - i
        return chi[i] + f * (chi[i + 1] - chi[i])


def comoving_distance(cosmology, a):
    """chi(a) in Mpc: pyccl when given a pyccl Cosmology, else the object's own method."""
    if hasattr(cosmology, "comoving_radial_distance"):
        return np.asarray(cosmology.comoving_radial_distance(a), dtype=np.float64)
    import pyccl as ccl  # a pyccl Cosmology (only reachable when pyccl is installed)
    return np.asarray(ccl.comoving_radial_distance(cosmology, a), dtype=np.float64)


_DEFAULT = []


def default_cosmology():
    """The reference's fallback cosmology (measure_w_lightcone.py:127-131)."""
    if not _DEFAULT:
        try:
            import pyccl as ccl
            _DEFAULT.append(ccl.Cosmology(Omega_c=0.225, Omega_b=0.045, sigma8=0.8, h=0.7, n_s=1.0))
        except ImportError:
            _DEFAULT.append(FlatLCDM())
    return _DEFAULT[0]


def rotated_ellipticity(e1, e2):
    """c2 = -e cos 2phi_axis, s2 = -e sin 2phi_axis from e1, e2, with e and
    phi_axis derived as the reference does (measure_w_lightcone.py:134-139);
    the counter's e+ = c2 cos 2phi_sep + s2 sin 2phi_sep is then the
    reference's -e cos 2(phi_axis - phi_sep)."""
    theta = 1.0 / 2 * np.arctan2(e2, e1)
    axis = np.array([np.cos(theta), np.sin(theta)])
    axis_len = np.sqrt(np.sum(axis ** 2, axis=0))
    axis = axis / axis_len
    e = np.sqrt(e1 ** 2 + e2 ** 2)
    phi_axis = np.arctan2(axis[1], axis[0])
    return -e * np.cos(2 * phi_axis), -e * np.sin(2 * phi_axis)


def _same_labels(a, b) -> bool:
    """Whether two (position labels, shape labels, P) triples are equal."""
    return a[2] == b[2] and np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def rp_cut_threshold(rp_cut) -> float:
    """Largest squared projected separation s with sqrt(s) <= rp_cut, so that
    the reference's test sqrt(s) > rp_cut (measure_m_lightcone.py:164) is s > threshold."""
    if rp_cut is None:
        rp_cut = 0.0
    rp_cut = float(rp_cut)
    if rp_cut < 0:
        return -1.0
    if rp_cut == 0:
        return 0.0
    big = rp_cut * rp_cut * 4 + 1.0
    first = binning._bisect_first_true(lambda s, t: np.sqrt(s) > rp_cut, 0.0, big, np.zeros(1))[0]
    return float(binning._prev(first))


class MeasureIALightcone(SimInfo):
    """w_gg, w_g+ and multipoles of lightcone data (measure_IA.py:265-1060).

    Same constructor and methods as the reference's MeasureIALightcone; the
    pair loops run on the GPU (there is no CPU fallback).
    """

    def __init__(self, data, randoms_data, output_file_name, separation_limits=[0.1, 20.0],
                 num_bins_r=8, num_bins_pi=20, pi_max=None, num_nodes=1, device=0):
        # MeasureIABase.__init__ with simulation=False, boxsize=None, periodicity=False
        SimInfo.__init__(self, False, None, None)
        self.data = data
        self.output_file_name = output_file_name
        self.periodicity = False
        try:
            self.Num_position = len(data["Position"])
            self.Num_shape = len(data["Position_shape_sample"])
        except (KeyError, TypeError):
            try:
                self.Num_position = len(data["RA"])
                self.Num_shape = len(data["RA_shape_sample"])
            except (KeyError, TypeError):
                self.Num_position = 0
                self.Num_shape = 0
                print("Warning: no Postion or Position_shape_sample given.")
        if self.Num_position > 0:
            if "weight" not in self.data:
                self.data["weight"] = np.ones(self.Num_position)
            if "weight_shape_sample" not in self.data:
                self.data["weight_shape_sample"] = np.ones(self.Num_shape)
        self.r_min = separation_limits[0]
        self.r_max = separation_limits[1]
        self.num_bins_r = num_bins_r
        self.num_bins_pi = num_bins_pi
        self.r_bins = np.logspace(np.log10(self.r_min), np.log10(self.r_max), self.num_bins_r + 1)
        if pi_max is None:
            if self.L_0p5 is None:
                raise ValueError("Both pi_max and boxsize are None. Provide input on one of them to "
                                 "determine the integration limit pi_max.")
            pi_max = self.L_0p5
        self.pi_bins = np.linspace(-pi_max, pi_max, self.num_bins_pi + 1)
        self.mu_r_bins = np.linspace(-1, 1, self.num_bins_pi + 1)
        self.num_nodes = num_nodes
        self.randoms_data = randoms_data
        self.data_dir = None
        self.num_samples = None
        self.device = device
        self._th = {}
        self._memo = None   # per-measurement cache of converted data columns
        self._share = None  # per-measurement plan of counts shared with the jackknife realisations
        self.device_counts = 0  # device counts made so far

    # shared post-processing of the box front-end (measure_IA_base.py:515-668)
    _store = MeasureIABox._store
    _measure_w_g_i = MeasureIABox._measure_w_g_i
    _measure_multipoles = MeasureIABox._measure_multipoles
    # covariance across datasets, inherited by the reference's lightcone class
    # from MeasureJackknife (measure_jackknife.py:485-648); both read only the
    # stored realisations
    measure_covariance_multiple_datasets = MeasureIABox.measure_covariance_multiple_datasets
    create_full_cov_matrix_projections = MeasureIABox.create_full_cov_matrix_projections

    # -- the reference's helpers --------------------------------------------------
    @staticmethod
    def get_ellipticity(e, phi):
        """e+ = e cos 2phi, ex = e sin 2phi (measure_IA_base.py:209-227)."""
        return e * np.cos(2 * phi), e * np.sin(2 * phi)

    # the rest of MeasureIABase's helpers, which the reference's lightcone
    # class inherits too (measure_IA_base.py:186-401)
    calculate_dot_product_arrays = staticmethod(MeasureIABox.calculate_dot_product_arrays)
    setdiff2D = staticmethod(MeasureIABox.setdiff2D)
    setdiff_omit = staticmethod(MeasureIABox.setdiff_omit)
    get_random_pairs = staticmethod(MeasureIABox.get_random_pairs)
    get_volume_spherical_cap = staticmethod(MeasureIABox.get_volume_spherical_cap)
    get_random_pairs_r_mur = MeasureIABox.get_random_pairs_r_mur

    def assign_jackknife_patches(self, data, randoms_data, num_jk):
        """k-means patches on (RA, DEC) of the randoms (measure_IA_base.py:745-800).

        The reference uses the third-party kmeans_radec package; it is used
        here the same way when installed, else this raises: pass jk_patches.
        """
        try:
            from kmeans_radec import kmeans_sample
        except ImportError as e:
            raise ImportError("kmeans_radec is not installed: pass jk_patches (patch numbers of the "
                              "position, shape and random samples) instead of num_jk") from e
        jk_patches = {}
        X = np.column_stack((randoms_data["RA"], randoms_data["DEC"]))
        km = kmeans_sample(X, num_jk, maxiter=100, tol=1.0e-5)
        jk_patches["randoms_position"] = km.labels
        jk_patches["randoms_shape"] = km.find_nearest(
            np.column_stack((randoms_data["RA_shape_sample"], randoms_data["DEC_shape_sample"])))
        jk_patches["position"] = km.find_nearest(np.column_stack((data["RA"], data["DEC"])))
        jk_patches["shape"] = km.find_nearest(
            np.column_stack((data["RA_shape_sample"], data["DEC_shape_sample"])))
        return jk_patches

    # -- device counts -------------------------------------------------------------
    def _engine(self):
        from .engine import get_engine
        return get_engine(self.device)

    def _thresholds(self, mode):
        th = self._th.get(mode)
        if th is None:
            if mode == "rppi":
                th = binning.rppi_thresholds(self.r_bins, self.pi_bins, self.r_min, self.r_max)
            else:
                th = binning.rmu_thresholds(self.r_bins, self.mu_r_bins, self.r_min, self.r_max)
            self._th[mode] = th
        return th

    def _select(self, masks, shapes):
        """The arrays a counter selects (measure_w_lightcone.py:82-117), with the
        reference's fill-in of missing weight masks (it mutates ``masks``)."""
        keys = ["Redshift", "Redshift_shape_sample", "RA", "RA_shape_sample", "DEC", "DEC_shape_sample"]
        if shapes:
            keys += ["e1", "e2"]
        if masks is None:
            out = {k: np.asarray(self.data[k]) for k in keys}
            out["weight"] = np.asarray(self.data["weight"])
            out["weight_shape_sample"] = np.asarray(self.data["weight_shape_sample"])
            return out
        out = {k: np.asarray(self.data[k])[masks[k]] for k in keys}
        if "weight" not in masks:
            masks["weight"] = np.ones(self.Num_position, dtype=bool)
        if "weight_shape_sample" not in masks:
            masks["weight_shape_sample"] = np.ones(self.Num_shape, dtype=bool)
        out["weight"] = np.asarray(self.data["weight"])[masks["weight"]]
        out["weight_shape_sample"] = np.asarray(self.data["weight_shape_sample"])[masks["weight_shape_sample"]]
        return out

    def _column(self, tag, keys, masks, fn):
        """fn of whole data columns (elementwise), then the selection of keys[0].

        Within one measurement the columns of self.data are converted once
        (comoving distances, cos dec, rotated ellipticities) and every count
        selects from the result: fn(x)[mask] equals fn(x[mask]) elementwise.
        Masks that differ between the keys are applied first instead."""
        full = [np.asarray(self.data[k]) for k in keys]
        m0 = None if masks is None else masks[keys[0]]
        if masks is not None and any(not (masks[k] is m0 or np.array_equal(masks[k], m0)) for k in keys[1:]):
            return fn(*[f[masks[k]] for f, k in zip(full, keys)])
        memo = getattr(self, "_memo", None)
        key = (tag,) + tuple(id(f) for f in full)
        hit = memo.get(key) if memo is not None else None
        if hit is not None and all(a is b for a, b in zip(hit[0], full)):
            val = hit[1]
        else:
            val = fn(*full)
            if memo is not None:
                memo[key] = (full, val)
        if m0 is None:
            return val
        return tuple(v[m0] for v in val) if isinstance(val, tuple) else val[m0]

    def _arrays(self, sel, shapes, over_h, cosmology, print_num, rp_cut=None, patches=None, masks=None):
        """LightconeArrays of a selection (measure_w_lightcone.py:118-139)."""
        n_p, n_s = len(sel["RA"]), len(sel["RA_shape_sample"])
        if print_num:
            print(f"There are {n_s} galaxies in the shape sample and {n_p} galaxies in the position sample.")
        if cosmology is None:
            cosmology = default_cosmology()
            if print_num:
                print("No cosmology given, using Omega_m=0.27, Omega_b=0.045, sigma8=0.8, h=0.7, n_s=1.")
        h = cosmology["h"]

        def dist(z):
            chi = comoving_distance(cosmology, 1 / (1 + np.asarray(z, dtype=np.float64)))
            return chi * h if over_h else chi

        ctag = ("chi", id(cosmology), bool(over_h))
        chi = self._column(ctag, ["Redshift"], masks, dist)
        chi_s = self._column(ctag, ["Redshift_shape_sample"], masks, dist)
        dec = np.asarray(sel["DEC"], dtype=np.float64)
        cos_dec = self._column("cos", ["DEC"], masks,
                               lambda d: np.cos(np.asarray(d, dtype=np.float64) / 180 * np.pi))
        c2 = s2 = None
        if shapes:
            c2, s2 = self._column("ell", ["e1", "e2"], masks, lambda a, b: rotated_ellipticity(
                np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)))
        lc = LightconeArrays(
            ra=np.asarray(sel["RA"], dtype=np.float64), dec=dec, chi=chi,
            cos_dec=cos_dec,
            ra_s=np.asarray(sel["RA_shape_sample"], dtype=np.float64),
            dec_s=np.asarray(sel["DEC_shape_sample"], dtype=np.float64), chi_s=chi_s,
            weight=np.asarray(sel["weight"], dtype=np.float64),
            weight_shape=np.asarray(sel["weight_shape_sample"], dtype=np.float64),
            c2=c2, s2=s2, proj_scale=float(h) if over_h else 1.0, rp2_cut=rp_cut_threshold(rp_cut))
        if patches is not None:
            lc.patch, lc.patch_s, lc.n_patches = patches
        return lc

    def _count(self, mode, masks, shapes, over_h, cosmology, print_num, rp_cut=None, patches=None):
        share = getattr(self, "_share", None)
        key = None
        if share is not None and masks is None:
            ip, is_ = id(self.data["RA"]), id(self.data["RA_shape_sample"])
            key = (mode, bool(shapes), ip, is_, bool(over_h), id(cosmology), rp_cut)
            if patches is None and ip in share["pos"] and is_ in share["shape"]:
                # a measurement whose jackknife realisations come next: count
                # once with the realisations' patch tallies; its totals are
                # the measurement and the realisations re-use the count
                patches = self._patch_labels(share["pos"][ip], share["shape"][is_], None)[-1]
        if key is not None and patches is not None:
            hit = share["done"].get(key)
            if hit is not None and _same_labels(hit[0], patches):
                return hit[1]
        sel = self._select(masks, shapes)
        lc = self._arrays(sel, shapes, over_h, cosmology, print_num, rp_cut, patches, masks)
        res = self._engine().count_lightcone(lc, self._thresholds(mode))
        self.device_counts += 1
        if key is not None and patches is not None:
            share["done"][key] = (patches, res)
        return res

    def _centres(self, mode):
        sep = estimators.bin_centres(self.r_bins)
        second = self.pi_bins if mode == "rppi" else self.mu_r_bins
        return sep, estimators.bin_centres(second)

    # -- counters (write or return, as the reference's) --------------------------------
    def _write_measure(self, group, names, dataset_name, jk_group_name, SplusD, ScrossD, DD, sep, second):
        corr = SplusD / DD
        st, own = self._store(None)
        try:
            g = f"{self.snap_group}/{group}/xi_g_plus/{jk_group_name}"
            st.write(g, dataset_name, corr)
            st.write(g, dataset_name + "_SplusD", SplusD)
            st.write(g, dataset_name + f"_{names[0]}", sep)
            st.write(g, dataset_name + f"_{names[1]}", second)
            g = f"{self.snap_group}/{group}/xi_g_cross/{jk_group_name}"
            st.write(g, dataset_name + "_ScrossD", ScrossD)
            st.write(g, dataset_name + f"_{names[0]}", sep)
            st.write(g, dataset_name + f"_{names[1]}", second)
            g = f"{self.snap_group}/{group}/xi_gg/{jk_group_name}"
            st.write(g, dataset_name + "_DD", DD)
            st.write(g, dataset_name + f"_{names[0]}", sep)
            st.write(g, dataset_name + f"_{names[1]}", second)
        finally:
            if own:
                st.close()

    def _write_count(self, group, names, dataset_name, jk_group_name, data_suffix, DD, sep, second):
        st, own = self._store(None)
        try:
            g = f"{self.snap_group}/{group}/xi_gg/{jk_group_name}"
            st.write(g, dataset_name + data_suffix, DD)
            st.write(g, dataset_name + f"_{names[0]}", sep)
            st.write(g, dataset_name + f"_{names[1]}", second)
        finally:
            if own:
                st.close()

    @staticmethod
    def _fix_dd(DD):
        DD = np.array(DD, dtype=np.float64)
        DD[np.where(DD == 0)] = 1
        return DD

    def _measure(self, mode, dataset_name, masks, return_output, print_num, over_h, cosmology,
                 jk_group_name, rp_cut=None):
        res = self._count(mode, masks, True, over_h, cosmology, print_num, rp_cut)
        return self._emit_measure(mode, dataset_name, return_output, jk_group_name,
                                  res.splus[0], res.scross[0], res.dd[0])

    def _emit_measure(self, mode, dataset_name, return_output, jk_group_name, SplusD, ScrossD, DD):
        DD = self._fix_dd(DD)
        sep, second = self._centres(mode)
        group, names = ("w", ("rp", "pi")) if mode == "rppi" else ("multipoles", ("r", "mu_r"))
        if self.output_file_name is not None and not return_output:
            self._write_measure(group, names, dataset_name, jk_group_name, SplusD, ScrossD, DD, sep, second)
            return None
        return SplusD, DD, sep, second

    def _pairs(self, mode, dataset_name, masks, return_output, print_num, over_h, cosmology,
               data_suffix, jk_group_name, rp_cut=None):
        res = self._count(mode, masks, False, over_h, cosmology, print_num, rp_cut)
        return self._emit_count(mode, dataset_name, return_output, jk_group_name, data_suffix, res.dd[0])

    def _emit_count(self, mode, dataset_name, return_output, jk_group_name, data_suffix, DD):
        DD = self._fix_dd(DD)
        sep, second = self._centres(mode)
        group, names = ("w", ("rp", "pi")) if mode == "rppi" else ("multipoles", ("r", "mu_r"))
        if self.output_file_name is not None and not return_output:
            self._write_count(group, names, dataset_name, jk_group_name, data_suffix, DD, sep, second)
            return None
        return DD, sep, second

    def _measure_xi_rp_pi_lightcone_brute(self, dataset_name, masks=None, return_output=False,
                                          print_num=True, over_h=False, cosmology=None, jk_group_name=""):
        """S+D, SxD, DD of (r_p, pi) (measure_w_lightcone.py:45-214)."""
        return self._measure("rppi", dataset_name, masks, return_output, print_num, over_h, cosmology,
                             jk_group_name)

    def _count_pairs_xi_rp_pi_lightcone_brute(self, dataset_name, masks=None, return_output=False,
                                              print_num=True, over_h=False, cosmology=None,
                                              data_suffix="_DD", jk_group_name=""):
        """DD of (r_p, pi) (measure_w_lightcone.py:216-343)."""
        return self._pairs("rppi", dataset_name, masks, return_output, print_num, over_h, cosmology,
                           data_suffix, jk_group_name)

    def _measure_xi_r_mur_lightcone_brute(self, dataset_name, masks=None, return_output=False,
                                          print_num=True, over_h=True, cosmology=None, rp_cut=None,
                                          jk_group_name=""):
        """S+D, SxD, DD of (r, mu_r) (measure_m_lightcone.py:45-219; over_h defaults to True there)."""
        return self._measure("rmu", dataset_name, masks, return_output, print_num, over_h, cosmology,
                             jk_group_name, rp_cut)

    def _count_pairs_xi_r_mur_lightcone_brute(self, dataset_name, masks=None, return_output=False,
                                              print_num=True, over_h=False, cosmology=None, rp_cut=None,
                                              data_suffix="_DD", jk_group_name=""):
        """DD of (r, mu_r) (measure_m_lightcone.py:221-360)."""
        return self._pairs("rmu", dataset_name, masks, return_output, print_num, over_h, cosmology,
                           data_suffix, jk_group_name, rp_cut)

    # -- jackknife realisations (measure_jackknife.py:59-483) ----------------------------
    def _patch_labels(self, patches_pos, patches_shape, masks, warn=False):
        """Patch labels of one realisation count: (min, max, number of patches,
        position and shape selections, (position labels, shape labels, P)).
        Labels run 0..num_patches-1; a shape label outside the positions' range
        gets the extra tally num_patches and is never removed."""
        patches_pos = np.asarray(patches_pos)
        patches_shape = np.asarray(patches_shape)
        min_patch, max_patch = int(np.min(patches_pos)), int(np.max(patches_pos))
        num_patches = max_patch - min_patch + 1
        if warn and np.min(patches_shape) != min_patch:
            print("Warning! Minimum patch number of shape sample is not equal to minimum patch number of "
                  "position sample.")
        if warn and np.max(patches_shape) != max_patch:
            print("Warning! Maximum patch number of shape sample is not equal to maximum patch number of "
                  "position sample.")
        mp = np.ones(len(patches_pos), dtype=bool)
        ms = np.ones(len(patches_shape), dtype=bool)
        if masks is not None:
            # the reference's (patches != i) * masks["Redshift"] (raises like it on a length mismatch)
            mp = np.asarray(mp * masks["Redshift"], dtype=bool)
            ms = np.asarray(ms * masks["Redshift_shape_sample"], dtype=bool)
        lp = patches_pos[mp] - min_patch
        ls = patches_shape[ms] - min_patch
        P = num_patches
        outside = (ls < 0) | (ls >= num_patches)
        if outside.any():
            ls = np.where(outside, num_patches, ls)
            P = num_patches + 1
        return min_patch, max_patch, num_patches, mp, ms, (lp.astype(np.int32), ls.astype(np.int32), P)

    def _realisation_counts(self, patches_pos, patches_shape, masks, mode, shapes, over_h, cosmology,
                            rp_cut, num_sample_names):
        """One device count of the patch-labelled sample; returns (patch numbers,
        per-realisation (S+D, SxD, DD)).  Realisation i keeps the objects with
        patch != i and, with masks, masks["Redshift"] / ["Redshift_shape_sample"]
        (measure_jackknife.py:84-110); it is the total minus tally[i]."""
        min_patch, max_patch, num_patches, mp, ms, patches = self._patch_labels(patches_pos, patches_shape,
                                                                                 masks, warn=True)
        # sum(mask) of every realisation (measure_jackknife.py:89-90): the
        # selected objects minus those of patch i, counted in one pass
        def per_patch(labels, sel):
            lab = labels[sel]
            inside = (lab >= min_patch) & (lab <= max_patch)
            return lab.size, np.bincount((lab[inside] - min_patch).astype(np.int64), minlength=num_patches)

        n_s, c_s = per_patch(np.asarray(patches_shape), ms)
        n_p, c_p = per_patch(np.asarray(patches_pos), mp)
        for i in range(min_patch, max_patch + 1):
            self.num_samples[f"{i}"][num_sample_names[0]] = int(n_s - c_s[i - min_patch])
            self.num_samples[f"{i}"][num_sample_names[1]] = int(n_p - c_p[i - min_patch])
        keys = {"Redshift": mp, "Redshift_shape_sample": ms, "RA": mp, "RA_shape_sample": ms, "DEC": mp,
                "DEC_shape_sample": ms, "weight": mp, "weight_shape_sample": ms}
        if shapes:
            keys.update({"e1": ms, "e2": ms})
        # without masks the selection is everything: the count is the one a
        # measurement made with the same patch labels, when there was one
        res = self._count(mode, None if masks is None else keys, shapes, over_h, cosmology, False, rp_cut,
                          patches=patches)
        reals = {}
        for i in range(num_patches):
            k = 1 + i
            sp = res.splus[0] - res.splus[k] if shapes else None
            sx = res.scross[0] - res.scross[k] if shapes else None
            reals[min_patch + i] = (sp, sx, res.dd[0] - res.dd[k])
        return min_patch, max_patch, reals

    def _measure_jackknife_realisations_lightcone(self, patches_pos, patches_shape, masks=None,
                                                  corr_type=["both", "w"], dataset_name="", rp_cut=None,
                                                  over_h=False, cosmology=None, count_pairs=False,
                                                  data_suffix="", num_sample_names=["S", "D"]):
        """Realisations written as the single-process loop writes them
        (measure_jackknife.py:59-171): each patch's counter output under
        {dataset_name}_jk{num_patches}."""
        if count_pairs and data_suffix == "":
            raise ValueError("Enter a data suffix (like _DD) for your pair count.")
        mode = "rmu" if corr_type[1] == "multipoles" else "rppi"
        mn, mx, reals = self._realisation_counts(patches_pos, patches_shape, masks, mode, not count_pairs,
                                                 over_h, cosmology, rp_cut, num_sample_names)
        num_patches = mx - mn + 1
        print(f"Calculating jackknife realisations for {num_patches} patches for {dataset_name}.")
        jk = f"{dataset_name}_jk{num_patches}"
        for i in range(mn, mx + 1):
            sp, sx, dd = reals[i]
            if count_pairs:
                self._emit_count(mode, f"{dataset_name}_{i}", False, jk, data_suffix, dd)
            else:
                self._emit_measure(mode, f"{dataset_name}_{i}", False, jk, sp, sx, dd)

    def _measure_jackknife_realisations_lightcone_multiprocessing(
            self, patches_pos, patches_shape, corr_type, dataset_name, masks=None, rp_cut=None,
            over_h=False, num_nodes=4, cosmology=None, count_pairs=False, data_suffix="",
            num_sample_names=["S", "D"]):
        """Realisations as the pooled loop writes them (measure_jackknife.py:265-483);
        num_nodes == 1 takes the single-process layout, like the reference."""
        if num_nodes == 1:
            self._measure_jackknife_realisations_lightcone(patches_pos, patches_shape, masks, corr_type,
                                                           dataset_name, rp_cut, over_h, cosmology,
                                                           count_pairs, data_suffix, num_sample_names)
            return
        if count_pairs is False:
            corr_type[0] = "both"
        if corr_type[0] == "both":
            data = [corr_type[1] + "_g_plus", corr_type[1] + "_gg"]
            corr_type_suff = ["_g_plus", "_gg"]
            xi_suff = ["SplusD", "DD"]
        elif corr_type[0] == "g+":
            data = [corr_type[1] + "_g_plus"]
            corr_type_suff = ["_g_plus"]
            xi_suff = ["SplusD"]
        elif corr_type[0] == "gg":
            data = [corr_type[1] + "_gg"]
            corr_type_suff = ["_gg"]
            xi_suff = ["DD"]
        else:
            raise KeyError("Unknown value for first entry of corr_type. Choose from [g+, gg, both]")
        if corr_type[1] == "multipoles":
            bin_names = ["r", "mu_r"]
        elif corr_type[1] == "w":
            bin_names = ["rp", "pi"]
        else:
            raise KeyError("Unknown value for second entry of corr_type. Choose from [multipoles, w_g_plus]")
        mode = "rmu" if corr_type[1] == "multipoles" else "rppi"
        mn, mx, reals = self._realisation_counts(patches_pos, patches_shape, masks, mode, not count_pairs,
                                                 over_h, cosmology, rp_cut, num_sample_names)
        num_patches = mx - mn + 1
        sep, second = self._centres(mode)
        st, own = self._store(None)
        try:
            for i in range(mn, mx + 1):
                sp, sx, dd = reals[i]
                dd = self._fix_dd(dd)
                if count_pairs:
                    result = (dd, sep, second)
                    for j, _ in enumerate(data):
                        g = f"{self.snap_group}/{corr_type[1]}/xi_gg/{dataset_name}_jk{num_patches}"
                        st.write(g, f"{dataset_name}_{i}{data_suffix}", result[j])
                        st.write(g, f"{dataset_name}_{i}_{bin_names[0]}", result[1])
                        st.write(g, f"{dataset_name}_{i}_{bin_names[1]}", result[2])
                else:
                    result = (sp, dd, sep, second)
                    for j, _ in enumerate(data):
                        g = (f"{self.snap_group}/{corr_type[1]}/xi{corr_type_suff[j]}/"
                             f"{dataset_name}_jk{num_patches}")
                        st.write(g, f"{dataset_name}_{i}_{xi_suff[j]}", result[j])
                        st.write(g, f"{dataset_name}_{i}_{bin_names[0]}", result[2])
                        st.write(g, f"{dataset_name}_{i}_{bin_names[1]}", result[3])
        finally:
            if own:
                st.close()

    def _obs_estimator(self, corr_type, IA_estimator, dataset_name, dataset_name_randoms, num_samples,
                       jk_group_name="", jk_group_name_randoms=""):
        """Cluster / galaxy estimators of one realisation from its stored counts
        (measure_IA_base.py:670-743); the formulas are estimators.LC_ESTIMATORS."""
        if IA_estimator not in ("clusters", "galaxies"):
            raise ValueError("Unknown input for IA_estimator, choose from [clusters, galaxies].")
        parts = {"both": ("g+", "gg"), "g+": ("g+",), "gg": ("gg",)}.get(corr_type[0], ())
        base = f"{self.snap_group}/{corr_type[1]}"
        g_plus = f"{base}/xi_g_plus/{jk_group_name}"
        gg = f"{base}/xi_gg/{jk_group_name}"
        # (group, dataset) of every named count; the random-shape count SR has its
        # own dataset only in a gg-only run, else it is the randoms' DD
        where = {
            "S+D": (g_plus, f"{dataset_name}_SplusD"),
            "S+R": (f"{base}/xi_g_plus/{jk_group_name_randoms}", f"{dataset_name_randoms}_SplusD"),
            "DD": (gg, f"{dataset_name}_DD"),
            "RD": (gg, f"{dataset_name}_RD"),
            "RR": (gg, f"{dataset_name}_RR"),
            "SR": ((gg, f"{dataset_name}_SR") if corr_type[0] == "gg"
                   else (f"{base}/xi_gg/{jk_group_name_randoms}", f"{dataset_name_randoms}_DD")),
        }
        target = {"g+": g_plus, "gg": gg}
        st, own = self._store(None)
        try:
            for part in parts:
                xi = estimators.lightcone_xi(IA_estimator, part,
                                             lambda name: st.read("/".join(where[name])), num_samples)
                st.write(target[part], dataset_name, xi)
        finally:
            if own:
                st.close()

    def _measure_jackknife_covariance_lightcone(self, IA_estimator, corr_type, dataset_name, max_patch,
                                                min_patch=1, randoms_suf="_randoms"):
        """Every leave-one-patch-out realisation's estimators and projections, then
        mean / std / covariance of each projected correlation (measure_jackknife.py:172-263)."""
        suffixes = {"both": ("_g_plus", "_gg"), "g+": ("_g_plus",), "gg": ("_gg",)}.get(corr_type[0])
        if suffixes is None:
            raise KeyError("Unknown value for corr_type. Choose from [g+, gg, both]")
        patches = range(min_patch, max_patch + 1)
        n = len(patches)
        print(f"Calculating jackknife errors for {n} patches for {dataset_name} with "
              f"{dataset_name}{randoms_suf} as randoms.")
        jk = f"{dataset_name}_jk{n}"
        project = self._measure_w_g_i if corr_type[1] == "w" else self._measure_multipoles
        for b in patches:
            self._obs_estimator(corr_type, IA_estimator, f"{dataset_name}_{b}", f"{dataset_name}{randoms_suf}_{b}",
                                self.num_samples[f"{b}"], jk_group_name=jk,
                                jk_group_name_randoms=f"{dataset_name}{randoms_suf}_jk{n}")
            project(corr_type=corr_type[0], dataset_name=f"{dataset_name}_{b}", jk_group_name=jk)
        covs, stds = [], []
        st, own = self._store(None)
        try:
            for suffix in suffixes:
                group = f"{self.snap_group}/{corr_type[1]}{suffix}"
                reals = [st.read(f"{group}/{jk}/{dataset_name}_{b}") for b in patches]
                mean, std, cov = estimators.jackknife_combine(reals, self.num_bins_r)
                if self.output_file_name is None:
                    covs.append(cov)
                    stds.append(std)
                    continue
                for tag, value in (("mean", mean), ("jackknife", std), ("jackknife_cov", cov)):
                    st.write(group, f"{dataset_name}_{tag}_{n}", value)
        finally:
            if own:
                st.close()
        if self.output_file_name is None:
            return covs, stds

    # -- measurements (measure_IA.py:337-1060) --------------------------------------------
    def _setup(self, IA_estimator, masks, masks_randoms, jk_patches, num_jk, errors):
        if IA_estimator == "clusters":
            if self.randoms_data is None:
                print("No randoms given, correlation defined as S+D/DD")
                raise KeyError("This version does not work yet, add randoms.")
            print("xi_g+ defined as S+D/SD - S+R/SR, xi_gg as (SD - RD - SR)/RR - 1")
            if masks is not None and masks_randoms is None:
                print("Warning, masks given for data vector but not for randoms.")
        elif IA_estimator == "galaxies":
            if self.randoms_data is None:
                raise KeyError("No randoms given. Please provide input.")
            print("xi_g+ defined as (S+D - S+R)/RR, xi_gg as (SD - RD - SR)/RR - 1")
            if masks is not None and masks_randoms is None:
                print("Warning, masks given for data vector but not for randoms.")
            print("WARNING: this version of the code has not been fully validated. Proceed with caution.")
        else:
            raise KeyError("Unknown input for IA_estimator, choose from [clusters, galaxies].")
        data = self.data
        R = self.randoms_data
        one_random_sample = "RA_shape_sample" not in R
        if one_random_sample:
            R["RA_shape_sample"] = R["RA"]
            R["DEC_shape_sample"] = R["DEC"]
            R["Redshift_shape_sample"] = R["Redshift"]
        if "weight" not in R:
            R["weight"] = np.ones(len(R["RA"]))
        if "weight_shape_sample" not in R:
            R["weight_shape_sample"] = R["weight"] if one_random_sample else np.ones(len(R["RA_shape_sample"]))
        if errors:
            if jk_patches is None:
                if num_jk is not None:
                    jk_patches = self.assign_jackknife_patches(data, R, num_jk)
                else:
                    raise ValueError("Set calc_errors to False, or provide either jk_patches or num_jk input.")
            elif one_random_sample:
                jk_patches["randoms_position"] = jk_patches["randoms"]
                jk_patches["randoms_shape"] = jk_patches["randoms"]
        self.data_dir = data
        if "weight" not in self.data_dir:
            self.data_dir["weight"] = np.ones(len(self.data_dir["RA"]))
        if "weight_shape_sample" not in self.data_dir:
            self.data_dir["weight_shape_sample"] = np.ones(len(self.data_dir["RA_shape_sample"]))
        ns = {}
        if masks is None:
            ns["D"] = len(self.data_dir["RA"])
            ns["S"] = len(self.data_dir["RA_shape_sample"])
        else:
            ns["D"] = len(np.asarray(self.data_dir["RA"])[masks["RA"]])
            ns["S"] = len(np.asarray(self.data_dir["RA_shape_sample"])[masks["RA_shape_sample"]])
        if masks_randoms is None:
            ns["R_D"] = len(R["RA"])
            ns["R_S"] = len(R["RA_shape_sample"])
        else:
            ns["R_D"] = len(np.asarray(R["RA"])[masks_randoms["RA"]])
            ns["R_S"] = len(np.asarray(R["RA_shape_sample"])[masks_randoms["RA_shape_sample"]])
        return data, jk_patches, ns

    def _combo(self, pos, shp, shapes):
        """Data dict of one position/shape combination (measure_IA.py:461-545)."""
        D, R = self.data_dir, self.randoms_data
        P = D if pos == "D" else R
        S = D if shp == "D" else R
        out = {"Redshift": P["Redshift"], "Redshift_shape_sample": S["Redshift_shape_sample"],
               "RA": P["RA"], "RA_shape_sample": S["RA_shape_sample"], "DEC": P["DEC"],
               "DEC_shape_sample": S["DEC_shape_sample"]}
        if shapes:
            out["e1"] = S["e1"]
            out["e2"] = S["e2"]
        out["weight"] = P["weight"]
        out["weight_shape_sample"] = S["weight_shape_sample"]
        return out

    def _share_plan(self, jk_patches):
        """Which patch labels belong to which position / shape arrays, so that
        each measurement count (S+D, S+R, SD, SR, RD, RR) is made once with
        the patch tallies its jackknife realisations need.  Off when a data
        array is also a randoms array (the labels would be ambiguous)."""
        D, R = self.data_dir, self.randoms_data
        pos = {id(D["RA"]): jk_patches["position"], id(R["RA"]): jk_patches["randoms_position"]}
        shape = {id(D["RA_shape_sample"]): jk_patches["shape"],
                 id(R["RA_shape_sample"]): jk_patches["randoms_shape"]}
        if len(pos) < 2 or len(shape) < 2:
            return None
        return {"pos": pos, "shape": shape, "done": {}}

    def _run(self, group, IA_estimator, dataset_name, corr_type, jk_patches, num_jk, errors, masks,
             masks_randoms, cosmology, over_h, rp_cut):
        mode = "rppi" if group == "w" else "rmu"
        self._memo = {}
        data, jk_patches, num_samples = self._setup(IA_estimator, masks, masks_randoms, jk_patches, num_jk,
                                                    errors)
        self._share = self._share_plan(jk_patches) if errors and masks is None else None
        if mode == "rppi":
            measure = lambda **k: self._measure_xi_rp_pi_lightcone_brute(**k)  # noqa: E731
            pairs = lambda **k: self._count_pairs_xi_rp_pi_lightcone_brute(**k)  # noqa: E731
            extra = {}
        else:
            measure = lambda **k: self._measure_xi_r_mur_lightcone_brute(**k)  # noqa: E731
            pairs = lambda **k: self._count_pairs_xi_r_mur_lightcone_brute(**k)  # noqa: E731
            extra = {"rp_cut": rp_cut}
        try:
            # S+D and S+R
            if corr_type == "g+" or corr_type == "both":
                self.data = self.data_dir
                measure(masks=masks, dataset_name=dataset_name, over_h=over_h, cosmology=cosmology, **extra)
                self.data = self._combo("R", "D", True)
                measure(masks=masks, dataset_name=f"{dataset_name}_randoms", over_h=over_h,
                        cosmology=cosmology, **extra)
            # SD, SR (gg only: 'both' has them from the S+D / S+R calls), RD, RR
            if corr_type == "gg":
                self.data = self._combo("D", "D", False)
                pairs(masks=masks, dataset_name=dataset_name, over_h=over_h, cosmology=cosmology,
                      data_suffix="_DD", **extra)
                self.data = self._combo("R", "D", False)
                pairs(masks=masks, dataset_name=dataset_name, over_h=over_h, cosmology=cosmology,
                      data_suffix="_SR", **extra)
            if corr_type == "gg" or corr_type == "both":
                self.data = self._combo("D", "R", False)
                pairs(masks=masks, dataset_name=dataset_name, over_h=over_h, cosmology=cosmology,
                      data_suffix="_RD", **extra)
            if IA_estimator == "galaxies" or corr_type == "gg" or corr_type == "both":
                self.data = self._combo("R", "R", False)
                pairs(masks=masks, dataset_name=dataset_name, over_h=over_h, cosmology=cosmology,
                      data_suffix="_RR", **extra)
            self._obs_estimator([corr_type, group], IA_estimator, dataset_name, f"{dataset_name}_randoms",
                                num_samples)
            if group == "w":
                self._measure_w_g_i(corr_type=corr_type, dataset_name=dataset_name, return_output=False)
            else:
                self._measure_multipoles(corr_type=corr_type, dataset_name=dataset_name, return_output=False)
            if errors:
                self._errors(group, IA_estimator, dataset_name, corr_type, jk_patches, masks, cosmology,
                             over_h, rp_cut)
        finally:
            self.data = data
            self._memo = None
            self._share = None

    def _errors(self, group, IA_estimator, dataset_name, corr_type, jk_patches, masks, cosmology, over_h,
                rp_cut):
        self.num_samples = {}
        mn, mx = int(np.min(jk_patches["shape"])), int(np.max(jk_patches["shape"]))
        for n in range(mn, mx + 1):
            self.num_samples[f"{n}"] = {}
        real = self._measure_jackknife_realisations_lightcone_multiprocessing
        kw = dict(num_nodes=self.num_nodes, over_h=over_h, cosmology=cosmology)
        if group == "multipoles":
            kw["rp_cut"] = rp_cut
        if corr_type == "g+" or corr_type == "both":
            self.data = self.data_dir
            real(patches_pos=jk_patches["position"], patches_shape=jk_patches["shape"],
                 corr_type=[corr_type, group], masks=masks, dataset_name=dataset_name, count_pairs=False,
                 num_sample_names=["S", "D"], **kw)
            self.data = self._combo("R", "D", True)
            real(patches_pos=jk_patches["randoms_position"], patches_shape=jk_patches["shape"],
                 corr_type=[corr_type, group], masks=masks, dataset_name=f"{dataset_name}_randoms",
                 count_pairs=False, num_sample_names=["S", "R_D"], **kw)
        if corr_type == "gg":
            self.data = self._combo("D", "D", False)
            real(patches_pos=jk_patches["position"], patches_shape=jk_patches["shape"],
                 corr_type=["gg", group], dataset_name=dataset_name, count_pairs=True, data_suffix="_DD",
                 num_sample_names=["S", "D"], **kw)
            self.data = self._combo("R", "D", False)
            real(patches_pos=jk_patches["randoms_position"], patches_shape=jk_patches["shape"],
                 corr_type=["gg", group], dataset_name=dataset_name, count_pairs=True, data_suffix="_SR",
                 num_sample_names=["S", "R_D"], **kw)
        if corr_type == "gg" or corr_type == "both":
            self.data = self._combo("D", "R", False)
            real(patches_pos=jk_patches["position"], patches_shape=jk_patches["randoms_shape"],
                 corr_type=["gg", group], dataset_name=dataset_name, count_pairs=True, data_suffix="_RD",
                 num_sample_names=["R_S", "D"], **kw)
        if IA_estimator == "galaxies" or corr_type == "gg" or corr_type == "both":
            self.data = self._combo("R", "R", False)
            real(patches_pos=jk_patches["randoms_position"], patches_shape=jk_patches["randoms_shape"],
                 corr_type=["gg", group], dataset_name=dataset_name, count_pairs=True, data_suffix="_RR",
                 num_sample_names=["R_S", "R_D"], **kw)
        self._measure_jackknife_covariance_lightcone(IA_estimator=IA_estimator,
                                                     max_patch=int(np.max(jk_patches["shape"])),
                                                     min_patch=int(np.min(jk_patches["shape"])),
                                                     corr_type=[corr_type, group], dataset_name=dataset_name,
                                                     randoms_suf="_randoms")

    def measure_xi_w(self, IA_estimator, dataset_name, corr_type, jk_patches=None, num_jk=None,
                     measure_cov=True, masks=None, masks_randoms=None, cosmology=None, over_h=False):
        """xi_gg, xi_g+, w_gg, w_g+ and their jackknife covariance (measure_IA.py:337-690)."""
        self._run("w", IA_estimator, dataset_name, corr_type, jk_patches, num_jk, measure_cov, masks,
                  masks_randoms, cosmology, over_h, None)

    def measure_xi_multipoles(self, IA_estimator, dataset_name, corr_type, jk_patches=None, num_jk=None,
                              calc_errors=True, masks=None, masks_randoms=None, cosmology=None,
                              over_h=False, rp_cut=None):
        """xi(r, mu_r), its multipoles and their jackknife covariance (measure_IA.py:692-1060)."""
        self._run("multipoles", IA_estimator, dataset_name, corr_type, jk_patches, num_jk, calc_errors,
                  masks, masks_randoms, cosmology, over_h, rp_cut)
