"""Host FP64 post-processing: analytic randoms, estimators, projections, covariance.

Everything downstream of the pair counts is O(bins x regions) host work in
the reference too; this module restates it vectorised, keeping the
reference's order of operations where the fixtures pin values bit-for-bit
(RR, w, jackknife mean/std/cov).
"""
from __future__ import annotations

import math

import numpy as np


def bin_centres(edges: np.ndarray) -> np.ndarray:
    """edges[:-1] + |half width| (measure_w_box_jk.py:243-246)."""
    d = (edges[1:] - edges[:-1]) / 2.0
    return edges[:-1] + abs(d)


def rr_rppi(r_bins, pi_bins, L3, n_pos, n_shape) -> np.ndarray:
    """RR = N_p N_s pi (r1^2 - r0^2) |pi1 - pi0| / L3 per bin (measure_IA_base.py:268-269)."""
    c = n_pos * n_shape * np.pi
    dr2 = r_bins[1:] ** 2 - r_bins[:-1] ** 2
    dpi = np.abs(pi_bins[1:] - pi_bins[:-1])
    return c * dr2[:, None] * dpi[None, :] / L3


def cap_volume(mu, r):
    """pi/3 r^3 (2 + mu) (1 - mu)^2 (measure_IA_base.py:291)."""
    return np.pi / 3.0 * r ** 3 * (2 + mu) * (1 - mu) ** 2


def rr_rmu(r_bins, mu_bins, L3, n_pos, n_shape) -> np.ndarray:
    """|(N_p - 1) N_s [V(mu0,r1) - V(mu1,r1) - (V(mu0,r0) - V(mu1,r0))] / L3|.

    measure_IA_base.py:337-351; note the (N_p - 1) factor of the 'cross' case.
    """
    r1 = r_bins[1:][:, None]
    r0 = r_bins[:-1][:, None]
    m1 = mu_bins[1:][None, :]
    m0 = mu_bins[:-1][None, :]
    rr = ((n_pos - 1.0) * n_shape
          * (cap_volume(m0, r1) - cap_volume(m1, r1) - (cap_volume(m0, r0) - cap_volume(m1, r0)))
          / L3)
    return np.abs(rr)


def responsivity_jk(total_a, total_w, region_a, region_w) -> np.ndarray:
    """R over the shapes outside each region: (A - a_i) / (W - w_i)."""
    return (total_a - region_a) / (total_w - region_w)


def w_projection(xi, pi_bins, stored_pi=None) -> np.ndarray:
    """w = sum over pi of xi |dpi| (measure_IA_base.py:553-562).

    ``stored_pi`` are the pi centres written next to xi; like the reference,
    a mismatch with the object's own pi bins raises.
    """
    dpi = pi_bins[1:] - pi_bins[:-1]
    centres = pi_bins[:-1] + abs(dpi) / 2.0
    if stored_pi is not None and not sum(np.isin(stored_pi, centres)) == len(stored_pi):
        raise ValueError("Update pi bins in initialisation of object to match xi_g_plus dataset.")
    d = np.array([dpi] * len(xi[:, 0]))
    return np.sum(xi * abs(d), axis=1)


def legendre_assoc(l: int, m: int, mu):
    """P_l^m(mu) with the Condon-Shortley phase, for the (l, m) the estimators use.

    The reference evaluates scipy.special.lpmn(l, m, mu)[0][-1, -1]
    (measure_IA_base.py:642-643); only (0, 0) and (2, 2) occur there, and
    (2, 0) serves the optional gg quadrupole.
    """
    mu = np.asarray(mu, dtype=np.float64)
    if (l, m) == (0, 0):
        return np.ones_like(mu)
    if (l, m) == (2, 2):
        return 3.0 * (1.0 - mu * mu)
    if (l, m) == (2, 0):
        return 0.5 * (3.0 * mu * mu - 1.0)
    raise NotImplementedError((l, m))


def multipole(xi, mu_centres, mu_bins, l: int, m: int) -> np.ndarray:
    """(2l+1)/2 (l-m)!/(l+m)! sum_mu P_l^m(mu) xi dmu (measure_IA_base.py:628-655)."""
    nr = xi.shape[0]
    P = np.broadcast_to(legendre_assoc(l, m, mu_centres)[None, :], xi.shape)
    dmu = mu_bins[1:] - mu_bins[:-1]
    dmu = np.array(list(dmu) * nr).reshape((nr, len(dmu)))
    f = (2 * l + 1) / 2.0 * math.factorial(l - m) / math.factorial(l + m)
    return np.sum(f * P * xi * dmu, axis=1)


def jackknife_combine(realisations, num_bins_r):
    """Mean, std and covariance with the (n-1)/n rule (measure_IA_base.py:478-498)."""
    n = len(realisations)
    mean = np.zeros(num_bins_r)
    for x in realisations:
        mean += x
    mean /= n
    cov = np.zeros((num_bins_r, num_bins_r))
    std = np.zeros(num_bins_r)
    for x in realisations:
        d = x - mean
        std += d ** 2
        for i in np.arange(num_bins_r):
            cov[:, i] += d * (x[i] - mean[i])
    std *= (n - 1) / n
    std = np.sqrt(std)
    cov *= (n - 1) / n
    return mean, std, cov


class Estimates:
    """All grids of one counter call (the values the reference writes)."""

    def __init__(self, mode, r_bins, second_bins, boxsize, n_pos, n_shape, R, counts,
                 region_counts=None, region_R=None, num_box=0, variance=None):
        L3 = boxsize ** 3
        self.mode = mode
        rr_fn = rr_rppi if mode == "rppi" else rr_rmu
        self.RR = rr_fn(r_bins, second_bins, L3, n_pos, n_shape)
        self.DD = counts["DD"]
        self.SplusD = counts["SplusD"]
        self.ScrossD = counts["ScrossD"]
        self.xi_g_plus = self.SplusD / self.RR
        self.xi_g_cross = self.ScrossD / self.RR
        self.xi_gg = (self.DD / self.RR) - 1
        self.sigmasq = None if variance is None else variance / self.RR ** 2
        self.r_centres = bin_centres(r_bins)
        self.second_centres = bin_centres(second_bins)
        self.num_box = num_box
        self.R = R
        if num_box:
            volume_jk = L3 * (num_box - 1) / num_box
            npj = n_pos - region_counts[:, 0]
            nsj = n_shape - region_counts[:, 1]
            self.RR_jk = np.stack([rr_fn(r_bins, second_bins, volume_jk, int(a), int(b))
                                   for a, b in zip(npj, nsj)])
            self.R_jk = region_R
            num = self.SplusD * (2 * R) - counts["SplusD_jk"]
            self.real_xi_g_plus = num / (self.RR_jk * 2 * self.R_jk[:, None, None])
            self.real_SplusD = num / (2 * self.R_jk[:, None, None])
            self.real_DD = self.DD - counts["DD_jk"]
            self.real_xi_gg = (self.real_DD / self.RR_jk) - 1


# Lightcone estimators (measure_IA_base.py:670-743).  Each estimator is one
# formula over named pair counts -- S+D and DD of shapes x positions, S+R of
# shapes x random positions, SR, RD (random shapes x positions) and RR.  A
# count that involves a random catalogue is first brought to the size of the
# data catalogue it stands in for, by N_X / N_R_X for each random sample X it
# contains (S for shapes, D for positions), taken in the order listed.
LC_RANDOM_SAMPLES = {"S+D": (), "DD": (), "S+R": ("D",), "SR": ("D",), "RD": ("S",), "RR": ("S", "D")}

LC_ESTIMATORS = {
    ("clusters", "g+"): (("S+D", "DD", "S+R", "SR"), lambda c: c["S+D"] / c["DD"] - c["S+R"] / c["SR"]),
    ("clusters", "gg"): (("DD", "RD", "SR", "RR"), lambda c: (c["DD"] - c["RD"] - c["SR"]) / c["RR"] - 1),
    ("galaxies", "g+"): (("S+D", "S+R", "RR"), lambda c: (c["S+D"] - c["S+R"]) / c["RR"]),
    ("galaxies", "gg"): (("DD", "RD", "SR", "RR"), lambda c: (c["DD"] - c["RD"] - c["SR"]) / c["RR"] + 1),
}


def lightcone_xi(estimator, part, load, num_samples):
    """xi of one lightcone estimator ('clusters' / 'galaxies') for part 'g+' or
    'gg'; load(name) returns the stored count of that name, num_samples the
    catalogue sizes {"S", "D", "R_S", "R_D"}."""
    names, formula = LC_ESTIMATORS[(estimator, part)]
    counts = {}
    for name in names:
        x = load(name)
        ratios = [num_samples[s] / num_samples["R_" + s] for s in LC_RANDOM_SAMPLES[name]]
        if ratios:
            f = ratios[0]
            for r in ratios[1:]:
                f = f * r
            x = x * f
        counts[name] = x
    return formula(counts)
