"""MeasureIABox: the reference's box API with the pair loop on the GPU.

Drop-in for ``measureia.MeasureIABox`` (src/measureia/measure_IA.py:10-262):
same constructor, same ``measure_xi_w`` / ``measure_xi_multipoles``
signatures, same data-dict inputs and the same HDF5 output layout.  The
``backend`` switch chooses the device counter:

  "hip"        batched fast path (MIA_FAST) -- the default
  "hip-plain"  one-thread-per-shape FP64 global-atomic path (MIA_PLAIN)

(also settable with the MEASURE_IA_BACKEND environment variable).  There is no
CPU backend: without a GPU and the built library the count raises.

The reference picks one of three CPU counters from ``temp_file_path`` and
``num_nodes`` (measure_IA.py:120-160); they differ only in which datasets
they write.  The same choice here selects the layout:
  temp_file_path=False          -> brute layout (jackknife runs write _sigmasq = var/RR^2)
  a path or None, num_nodes == 1 -> tree layout (_sigmasq written as zeros)
  a path or None, num_nodes > 1  -> multiprocessing layout (no _sigmasq)
No temporary file is written: the data stay in device memory.  Unlike the
reference, temp_file_path=None is accepted (the GPU path needs no staging).
"""
from __future__ import annotations

import os
import threading
import time
from typing import Optional

import numpy as np

from . import binning, estimators, hostprep
from .hdf5io import H5Store
from .siminfo import SimInfo

BACKENDS = {"hip": "fast", "hip-plain": "plain"}


_HASH_CHUNK = 1 << 20


def _fingerprint(data) -> tuple:
    """Content key of a data dict: every array's dtype, shape and a 64-bit
    hash of its bytes (xxh3, or blake2b where xxhash is missing); scalars by
    value.  Two calls on unchanged data get the same key.

    The bytes are hashed in 1 MiB pieces with a yield of the interpreter lock
    after each (xxh3 holds it while it hashes), so a count running on another
    thread is not held up for the whole hash."""
    try:
        import xxhash

        def h(b):
            x = xxhash.xxh3_64()
            for i in range(0, len(b), _HASH_CHUNK):
                x.update(b[i:i + _HASH_CHUNK])
                time.sleep(0)
            return x.intdigest()
    except ImportError:  # pragma: no cover
        import hashlib

        def h(b):
            return hashlib.blake2b(b, digest_size=8).hexdigest()
    out = []
    for k in sorted(data):
        v = data[k]
        if isinstance(v, (np.ndarray, list, tuple)):
            a = np.ascontiguousarray(v)
            out.append((k, a.dtype.str, a.shape, h(memoryview(a).cast("B"))))
        else:
            out.append((k, repr(v)))
    return tuple(out)


def _cube_root(num_jk: int) -> int:
    """Integer cube root or ValueError (measure_IA.py:95-101)."""
    L = int(round(num_jk ** (1.0 / 3)))
    for c in (L - 1, L, L + 1):
        if c > 0 and c ** 3 == num_jk:
            return c
    raise ValueError(
        f"Use x^3 as input for num_jk, with x as an int. {float(int(num_jk ** (1. / 3)))},{num_jk ** (1. / 3)}")


def _corr_groups(corr_type: str, prefix: str):
    if corr_type == "both":
        return [f"{prefix}_g_plus", f"{prefix}_gg"]
    if corr_type == "g+":
        return [f"{prefix}_g_plus"]
    if corr_type == "gg":
        return [f"{prefix}_gg"]
    raise KeyError("Unknown value for corr_type. Choose from [g+, gg, both]")


class MeasureIABox(SimInfo):
    """w_gg, w_g+ and multipoles with jackknife covariance for periodic boxes."""

    def __init__(self, data, output_file_name, simulation=None, snapshot=None,
                 separation_limits=[0.1, 20.0], num_bins_r=8, num_bins_pi=20, pi_max=None,
                 boxsize=None, periodicity=True, num_nodes=1, *, backend: Optional[str] = None,
                 device: int = 0, tuning: Optional[dict] = None, verbose: bool = True):
        SimInfo.__init__(self, simulation, snapshot, boxsize)
        self.data = data
        self.output_file_name = output_file_name
        self.periodicity = periodicity
        self.num_nodes = num_nodes
        self.verbose = verbose
        self.backend = backend or os.environ.get("MEASURE_IA_BACKEND", "hip")
        if self.backend not in BACKENDS:
            raise ValueError(f"unknown backend {self.backend!r}; choose from {sorted(BACKENDS)}")
        self.device = device
        self.tuning = tuning
        self._cache = None  # host prep, device copy and device sorts of unchanged data
        # MeasureIABase.__init__ (measure_IA_base.py:135-167)
        try:
            self.Num_position = len(data["Position"])
            self.Num_shape = len(data["Position_shape_sample"])
        except (TypeError, KeyError):
            self.Num_position = 0
            self.Num_shape = 0
            if verbose:
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
        self.pi_max = pi_max
        self.pi_bins = np.linspace(-pi_max, pi_max, self.num_bins_pi + 1)
        self.mu_r_bins = np.linspace(-1, 1, self.num_bins_pi + 1)
        self._thresholds = {}
        self.last_timing = {}
        if verbose:
            print(f"MeasureIABox ({self.backend}): boxsize {self.boxsize}, "
                  f"{self.Num_shape} shapes, {self.Num_position} positions, r bins {self.r_bins}, "
                  f"{num_bins_pi} pi/mu bins, pi_max {pi_max}, output {self.output_file_name}")

    # -- reference static helpers (measure_IA_base.py:229-351) -----------------
    @staticmethod
    def calculate_dot_product_arrays(a1, a2):
        """Row-wise dot product, columns added in order (measure_IA_base.py:186-206)."""
        dot_product = np.zeros(np.shape(a1)[0])
        for i in range(np.shape(a1)[1]):
            dot_product += a1[:, i] * a2[:, i]
        return dot_product

    @staticmethod
    def get_ellipticity(e, phi):
        """e_+ = e cos 2phi, e_x = e sin 2phi (measure_IA_base.py:208-227)."""
        return e * np.cos(2 * phi), e * np.sin(2 * phi)

    @staticmethod
    def setdiff2D(a1, a2):
        """Row-wise np.setdiff1d of two nested lists (measure_IA_base.py:353-375)."""
        assert len(a1) == len(a2), "Lengths of lists where each row is to be compared, are not the same."
        return [np.setdiff1d(a1[i], a2[i]) for i in range(len(a1))]

    @staticmethod
    def setdiff_omit(a1, a2, incl_ind):
        """np.setdiff1d(a1[i], a2) for the rows i in incl_ind (measure_IA_base.py:377-401)."""
        return [np.setdiff1d(a1[i], a2) for i in range(len(a1)) if np.isin(i, incl_ind)]

    @staticmethod
    def get_random_pairs(rp_max, rp_min, pi_max, pi_min, L3, corrtype, Num_position, Num_shape):
        if corrtype == "auto":
            return (Num_position - 1.0) * Num_shape / 2.0 * np.pi * (rp_max ** 2 - rp_min ** 2) * abs(
                pi_max - pi_min) / L3
        if corrtype == "cross":
            return Num_position * Num_shape * np.pi * (rp_max ** 2 - rp_min ** 2) * abs(pi_max - pi_min) / L3
        raise ValueError("Unknown input for corrtype, choose from auto or cross.")

    @staticmethod
    def get_volume_spherical_cap(mur, r):
        return estimators.cap_volume(mur, r)

    def get_random_pairs_r_mur(self, r_max, r_min, mur_max, mur_min, L3, corrtype, Num_position,
                               Num_shape):
        v = (estimators.cap_volume(mur_min, r_max) - estimators.cap_volume(mur_max, r_max)
             - (estimators.cap_volume(mur_min, r_min) - estimators.cap_volume(mur_max, r_min)))
        if corrtype == "auto":
            return abs((Num_position - 1.0) / 2.0 * Num_shape * v / L3)
        if corrtype == "cross":
            return abs((Num_position - 1.0) * Num_shape * v / L3)
        raise ValueError("Unknown input for corrtype, choose from auto or cross.")

    # -- jackknife regions (measure_IA_base.py:404-452) -------------------------
    def _get_jackknife_region_indices(self, masks, L_subboxes):
        if masks is None:
            pos = self.data["Position"]
            shp = self.data["Position_shape_sample"]
        else:
            pos = self.data["Position"][masks["Position"]]
            shp = self.data["Position_shape_sample"][masks["Position_shape_sample"]]
        return (hostprep.jackknife_regions(pos, self.boxsize, L_subboxes),
                hostprep.jackknife_regions(shp, self.boxsize, L_subboxes))

    # -- device counting --------------------------------------------------------
    def _engine(self):
        from .engine import get_engine
        return get_engine(self.device)

    def _bins(self, mode):
        if mode not in self._thresholds:
            if mode == "rppi":
                self._thresholds[mode] = binning.rppi_thresholds(self.r_bins, self.pi_bins, self.r_min,
                                                                 self.r_max)
            else:
                self._thresholds[mode] = binning.rmu_thresholds(self.r_bins, self.mu_r_bins,
                                                                self.r_min, self.r_max)
        return self._thresholds[mode]

    def _count(self, mode, masks, L_subboxes, ellipticity, variance):
        """Prepare, count on the device and assemble every estimator grid.

        With the batched backend and no masks, the host preparation, the
        device copy and the device cell sort (mia_prepare_sample) are kept
        while the data dict's contents stay the same: a second call on the
        same data (measure_xi_w then measure_xi_multipoles, or another
        corr_type) re-uses them; a count whose sort was made for the same
        binning mode and jackknife skips the sort too."""
        from .engine import SampleArrays
        t0 = time.perf_counter()
        th = self._bins(mode)
        jk = L_subboxes or 0
        reuse = masks is None and BACKENDS[self.backend] == "fast"
        cache, key, hashing = None, None, None
        if reuse and self._cache is not None:
            key = (_fingerprint(self.data), ellipticity)
            if self._cache["key"] == key:
                cache = self._cache
        if cache is None:
            self._cache = None  # releases the previous device copy and sorts
            upload = None
            if reuse:
                # the positions (the largest array, used as given) go to the
                # device while the host prepares the shape columns
                eng, dev_pos = self._engine(), {}
                src = self.data["Position"]

                def send():
                    try:
                        dev_pos["positions"] = eng.to_device({"positions": src})["positions"]
                    except BaseException as exc:  # re-raised by the caller
                        dev_pos["error"] = exc

                upload = threading.Thread(target=send, daemon=True)
                upload.start()
            p = hostprep.prepare(self.data, masks, ellipticity)
            host = dict(positions=p.positions, shapes=p.shapes, axis=p.axis, e=p.e,
                        weight=p.weight, weight_shape=p.weight_shape)
            if reuse:
                rest = self._engine().to_device({k: v for k, v in host.items() if k != "positions"})
                upload.join()
                if "error" in dev_pos:
                    raise dev_pos["error"]
                host = dict(rest, positions=dev_pos["positions"])
                cache = self._cache = {"key": key, "p": p, "arrays": host, "sorted": {}}
                if key is None:
                    # data seen for the first time: fingerprint it in a thread
                    # while the device counts (the library call releases the
                    # GIL), so a single call does not pay for the hash
                    data = self.data

                    def fingerprint(entry=cache):
                        entry["key"] = (_fingerprint(data), ellipticity)

                    hashing = threading.Thread(target=fingerprint, daemon=True)
                    hashing.start()
        else:
            p, host = cache["p"], cache["arrays"]
        if self.verbose:
            print(f"There are {p.n_shape} galaxies in the shape sample and {p.n_pos} galaxies in the "
                  f"position sample.")
        sample = SampleArrays(**host, boxsize=self.boxsize, los=p.los, jk_per_side=jk,
                              jk_sub=(self.L_0p5 * 2.0 / jk) if jk else 0.0,
                              periodic=bool(self.periodicity))
        t1 = time.perf_counter()
        eng = self._engine()
        if cache is not None:
            skey = (mode, jk, bool(self.periodicity), self.boxsize, repr(self.tuning),
                    np.asarray(th.r2_edges).tobytes(), np.asarray(th.x2_edges).tobytes())
            handle = cache["sorted"].get(skey)
            sorted_before = handle is not None
            if handle is None:
                handle = cache["sorted"][skey] = eng.prepare(sample, th, tuning=self.tuning)
            res = eng.count_prepared([handle], th, variance=variance, tuning=self.tuning)[0]
            res.stats["reused_sort"] = sorted_before
        else:
            res = eng.count(sample, th, method=BACKENDS[self.backend], variance=variance, tuning=self.tuning)
        if hashing is not None:
            hashing.join()
        t2 = time.perf_counter()
        inv2R = 1.0 / (2 * p.R)
        counts = {"DD": res.dd[0], "SplusD": res.splus[0] * inv2R, "ScrossD": res.scross * inv2R}
        K = jk ** 3 if jk else 0
        region_counts = region_R = None
        if K:
            counts["DD_jk"] = res.dd[1:]
            counts["SplusD_jk"] = res.splus[1:]
            rs = res.region_stats
            region_counts = rs[:, :2]
            region_R = estimators.responsivity_jk(p.sum_wa, p.sum_w, rs[:, 3], rs[:, 2])
        var = None
        if variance and res.variance is not None:
            var = res.variance * inv2R * inv2R
        sb = self.pi_bins if mode == "rppi" else self.mu_r_bins
        est = estimators.Estimates(mode, self.r_bins, sb, self.boxsize, int(p.n_pos), int(p.n_shape),
                                   p.R, counts, region_counts, region_R, K, var)
        est.npairs = res.npairs
        t3 = time.perf_counter()
        self.last_timing = dict(prepare_s=t1 - t0, count_s=t2 - t1, estimate_s=t3 - t2,
                                engine=res.stats)
        return est

    # -- output ------------------------------------------------------------------
    def _write_counter(self, st: H5Store, grp, est, ds, jk_group, layout):
        """Datasets of one counter call, in the reference's layout.

        JK: measure_w_box_jk.py:248-287 (brute), 498-537 (tree), 816-853
        (multiprocessing); no JK: measure_w_box.py:208-227; multipoles:
        measure_m_box_jk.py:261-300 and measure_m_box.py:408-427.
        """
        sn = self.snap_group
        second = "pi" if grp == "w" else "mu_r"
        rname = "rp" if grp == "w" else "r"
        K = est.num_box
        if K:
            sig = None
            if layout == "brute":
                sig = est.sigmasq
            elif layout == "tree":
                sig = np.zeros_like(est.RR)
            g = f"{sn}/{grp}/xi_g_plus/"
            st.write(g, ds, est.xi_g_plus)
            st.write(g, ds + "_SplusD", est.SplusD)
            st.write(g, ds + "_RR_g_plus", est.RR)
            if sig is not None:
                st.write(g, ds + "_sigmasq", sig)
            st.write(g, ds + f"_{rname}", est.r_centres)
            st.write(g, ds + f"_{second}", est.second_centres)
            g = f"{sn}/{grp}/xi_g_plus/{jk_group}"
            for i in range(K):
                st.write(g, ds + f"_{i}", est.real_xi_g_plus[i])
                st.write(g, ds + f"_{i}_SplusD", est.real_SplusD[i])
                st.write(g, ds + f"_{i}_RR", est.RR_jk[i])
                st.write(g, ds + f"_{i}_{rname}", est.r_centres)
                st.write(g, ds + f"_{i}_{second}", est.second_centres)
            g = f"{sn}/{grp}/xi_g_cross/{jk_group}"
            st.write(g, ds + "_ScrossD", est.ScrossD)
            st.write(g, ds, est.xi_g_cross)
            st.write(g, ds + "_RR_g_cross", est.RR)
            if sig is not None:
                st.write(g, ds + "_sigmasq", sig)
            st.write(g, ds + f"_{rname}", est.r_centres)
            st.write(g, ds + f"_{second}", est.second_centres)
            g = f"{sn}/{grp}/xi_gg/"
            st.write(g, ds, est.xi_gg)
            st.write(g, ds + "_DD", est.DD)
            st.write(g, ds + "_RR_gg", est.RR)
            if sig is not None:
                st.write(g, ds + "_sigmasq", sig)
            st.write(g, ds + f"_{rname}", est.r_centres)
            st.write(g, ds + f"_{second}", est.second_centres)
            g = f"{sn}/{grp}/xi_gg/{jk_group}"
            for i in range(K):
                st.write(g, ds + f"_{i}", est.real_xi_gg[i])
                st.write(g, ds + f"_{i}_DD", est.real_DD[i])
                st.write(g, ds + f"_{i}_RR", est.RR_jk[i])
                st.write(g, ds + f"_{i}_{rname}", est.r_centres)
                st.write(g, ds + f"_{i}_{second}", est.second_centres)
        else:
            g = f"{sn}/{grp}/xi_g_plus/{jk_group}"
            st.write(g, ds, est.xi_g_plus)
            st.write(g, ds + "_SplusD", est.SplusD)
            st.write(g, ds + "_RR_g_plus", est.RR)
            st.write(g, ds + f"_{rname}", est.r_centres)
            st.write(g, ds + f"_{second}", est.second_centres)
            g = f"{sn}/{grp}/xi_g_cross/{jk_group}"
            st.write(g, ds + "_ScrossD", est.ScrossD)
            st.write(g, ds, est.xi_g_cross)
            st.write(g, ds + "_RR_g_cross", est.RR)
            st.write(g, ds + f"_{rname}", est.r_centres)
            st.write(g, ds + f"_{second}", est.second_centres)
            g = f"{sn}/{grp}/xi_gg/{jk_group}"
            st.write(g, ds, est.xi_gg)
            st.write(g, ds + "_DD", est.DD)
            st.write(g, ds + "_RR_gg", est.RR)
            st.write(g, ds + f"_{rname}", est.r_centres)
            st.write(g, ds + f"_{second}", est.second_centres)

    def _store(self, store):
        if store is not None:
            return store, False
        if self.output_file_name is None:
            raise ValueError("output_file_name is None: nothing to write to")
        return H5Store(self.output_file_name, "a"), True

    # -- post-processing (measure_IA_base.py:454-668) -----------------------------
    def _measure_w_g_i(self, dataset_name, corr_type="both", return_output=False, jk_group_name="",
                       store=None):
        """w_gg / w_g+ from stored xi grids (measure_IA_base.py:515-575)."""
        xi_data = ["xi_g_plus", "xi_gg"] if corr_type == "both" else (
            ["xi_g_plus"] if corr_type == "g+" else (["xi_gg"] if corr_type == "gg" else None))
        if xi_data is None:
            raise KeyError("Unknown value for corr_type. Choose from [g+, gg, both]")
        st, own = self._store(store)
        try:
            for xd in xi_data:
                base = f"{self.snap_group}/w/{xd}/{jk_group_name}"
                xi = st.read(f"{base}/{dataset_name}")
                pi = st.read(f"{base}/{dataset_name}_pi")
                rp = st.read(f"{base}/{dataset_name}_rp")
                w = estimators.w_projection(xi, self.pi_bins, pi)
                if return_output:
                    return np.array([rp, w]).transpose()
                g = f"{self.snap_group}/{xd.replace('xi', 'w')}/{jk_group_name}"
                st.write(g, dataset_name + "_rp", rp)
                st.write(g, dataset_name, w)
        finally:
            if own:
                st.close()

    def _measure_multipoles(self, dataset_name, corr_type="both", return_output=False,
                            jk_group_name="", store=None):
        """multipoles_gg (l=0) / multipoles_g_plus (l=m=2) (measure_IA_base.py:577-668)."""
        if corr_type == "g+":
            todo = [("g_plus", 2)]
        elif corr_type == "gg":
            todo = [("gg", 0)]
        elif corr_type == "both":
            todo = [("g_plus", 2), ("gg", 0)]
        else:
            raise KeyError("Unknown value for corr_type. Choose from [g+, gg, both]")
        st, own = self._store(store)
        try:
            out = []
            for name, l in todo:
                base = f"{self.snap_group}/multipoles/xi_{name}/{jk_group_name}"
                xi = st.read(f"{base}/{dataset_name}")
                mu = st.read(f"{base}/{dataset_name}_mu_r")
                m = estimators.multipole(xi, mu, self.mu_r_bins, l, l)
                sep = estimators.bin_centres(self.r_bins)
                if return_output:
                    out.append(np.array([sep, m]).transpose())
                    continue
                g = f"{self.snap_group}/multipoles_{name}/{jk_group_name}"
                st.write(g, dataset_name + "_r", sep)
                st.write(g, dataset_name, m)
            if return_output:
                return out
        finally:
            if own:
                st.close()

    def _combine_jackknife_information(self, dataset_name, jk_group_name, corr_group, num_box,
                                       return_output=False, store=None):
        """Mean, std, covariance of the realisations (measure_IA_base.py:454-513)."""
        st, own = self._store(store)
        covs, stds = [], []
        try:
            for cg in corr_group:
                base = f"{self.snap_group}/{cg}/{jk_group_name}"
                reals = [st.read(f"{base}/{dataset_name}_{b}") for b in range(num_box)]
                mean, std, cov = estimators.jackknife_combine(reals, self.num_bins_r)
                if return_output:
                    covs.append(cov)
                    stds.append(std)
                else:
                    g = f"{self.snap_group}/{cg}"
                    st.write(g, dataset_name + "_mean_" + str(num_box), mean)
                    st.write(g, dataset_name + "_jackknife_" + str(num_box), std)
                    st.write(g, dataset_name + "_jackknife_cov_" + str(num_box), cov)
        finally:
            if own:
                st.close()
        if return_output:
            return covs, stds

    def measure_gg_quadrupole(self, dataset_name, num_jk=0, return_output=False):
        """l = 2 multipole of xi_gg(r, mu_r), read from the stored grids.

        An extension: the reference writes only l = 0 for gg
        (measure_IA_base.py:577-668).  The same weights are applied with
        P_2(mu) = (3 mu^2 - 1)/2 at the stored mu centres:
            xi_gg,2(r) = (2l+1)/2 sum_mu P_2(mu_c) xi_gg(r, mu) dmu.
        Writes multipoles_gg_l2/{dataset_name} (and _r); with num_jk > 0 also
        every realisation under multipoles_gg_l2/{dataset_name}_jk{num_jk} and
        their mean, std and covariance with the (n-1)/n rule, as for l = 0.
        Call after measure_xi_multipoles wrote the grids.
        """
        st, own = self._store(None)
        try:
            def one(name, jk_group):
                base = f"{self.snap_group}/multipoles/xi_gg/{jk_group}"
                xi = st.read(f"{base}/{name}")
                mu = st.read(f"{base}/{name}_mu_r")
                return estimators.multipole(xi, mu, self.mu_r_bins, 2, 0)

            sep = estimators.bin_centres(self.r_bins)
            q = one(dataset_name, "")
            if return_output:
                return np.array([sep, q]).transpose()
            g = f"{self.snap_group}/multipoles_gg_l2/"
            st.write(g, dataset_name + "_r", sep)
            st.write(g, dataset_name, q)
            if num_jk > 0:
                jk = f"{dataset_name}_jk{num_jk}"
                reals = []
                for i in range(num_jk):
                    qi = one(f"{dataset_name}_{i}", jk)
                    st.write(g + jk, f"{dataset_name}_{i}_r", sep)
                    st.write(g + jk, f"{dataset_name}_{i}", qi)
                    reals.append(qi)
                mean, std, cov = estimators.jackknife_combine(reals, self.num_bins_r)
                st.write(g, dataset_name + "_mean_" + str(num_jk), mean)
                st.write(g, dataset_name + "_jackknife_" + str(num_jk), std)
                st.write(g, dataset_name + "_jackknife_cov_" + str(num_jk), cov)
        finally:
            if own:
                st.close()

    # -- jackknife covariance across datasets (measure_jackknife.py:485-648) -----------
    def measure_covariance_multiple_datasets(self, corr_type, dataset_names, num_box=27,
                                             return_output=False):
        """Covariance of one dataset's or two datasets' jackknife realisations.

        The reference defines this on its lightcone class; it only reads
        realisations from the output file, so it serves box outputs as well.
        """
        from . import jackknife
        return jackknife.measure_covariance_multiple_datasets(
            self.output_file_name, self.snap_group, corr_type, dataset_names, num_box, self.num_bins_r,
            return_output)

    def create_full_cov_matrix_projections(self, corr_type, dataset_names=["LOS_x", "LOS_y", "LOS_z"],
                                           num_box=27, return_output=False):
        """Combined covariance of three projections and of every pair of them."""
        from . import jackknife
        return jackknife.create_full_cov_matrix_projections(
            self.output_file_name, self.snap_group, corr_type, dataset_names, num_box, self.num_bins_r,
            return_output)

    # -- public API ----------------------------------------------------------------
    def _layout(self, temp_file_path):
        if temp_file_path is False:
            return "brute"
        return "multiprocessing" if self.num_nodes > 1 else "tree"

    def _run(self, grp, dataset_name, corr_type, num_jk, temp_file_path, masks, ellipticity):
        mode = "rppi" if grp == "w" else "rmu"
        prefix = "w" if grp == "w" else "multipoles"
        groups = _corr_groups(corr_type, prefix)  # fail before the count, same KeyError
        L = _cube_root(num_jk) if num_jk > 0 else 0
        if self.data is not None and "RA" in self.data:
            print("Given data is lightcone, use measure_xi_w_lightcone method instead.")
        layout = self._layout(temp_file_path)
        est = self._count(mode, masks, L, ellipticity, variance=(layout == "brute" and L > 0))
        t0 = time.perf_counter()
        jk_group = f"{dataset_name}_jk{num_jk}" if num_jk > 0 else ""
        post = self._measure_w_g_i if grp == "w" else self._measure_multipoles
        st, own = self._store(None)
        try:
            self._write_counter(st, "w" if grp == "w" else "multipoles", est, dataset_name, jk_group,
                                layout)
            post(corr_type=corr_type, dataset_name=dataset_name, store=st)
            if num_jk > 0:
                for i in range(num_jk):
                    post(corr_type=corr_type, dataset_name=f"{dataset_name}_{i}",
                         jk_group_name=jk_group, store=st)
                self._combine_jackknife_information(dataset_name, jk_group, groups, num_jk, store=st)
        finally:
            if own:
                st.close()
        self.last_timing["write_s"] = time.perf_counter() - t0
        self.last_estimates = est

    def measure_xi_w(self, dataset_name, corr_type, num_jk=0, temp_file_path=None, masks=None,
                     ellipticity='distortion', chunk_size=1000):
        """xi_gg, xi_g+ on (r_p, pi) and w_gg, w_g+ with jackknife covariance (measure_IA.py:68-163)."""
        self._run("w", dataset_name, corr_type, num_jk, temp_file_path, masks, ellipticity)

    def measure_xi_multipoles(self, dataset_name, corr_type, num_jk=0, temp_file_path=None, masks=None,
                              rp_cut=None, ellipticity='distortion', chunk_size=1000):
        """xi on (r, mu_r) and the gg / g+ multipoles (measure_IA.py:165-262).

        rp_cut is accepted and ignored, as in the reference (it is never
        forwarded to the counters, measure_IA.py:218-259).
        """
        self._run("m", dataset_name, corr_type, num_jk, temp_file_path, masks, ellipticity)
