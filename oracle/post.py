"""Analytic randoms, estimators, projections and jackknife combination (oracle).

Test infrastructure only: imported by tests/, smoke() and bench.py's
cpu_baseline, never by the product package.
"""
import math

import numpy as np


def random_pairs_rppi(rp_max, rp_min, pi_max, pi_min, L3, Num_position, Num_shape):
    """'cross' RR of an (r_p, pi) bin (measure_IA_base.py:229-272)."""
    return Num_position * Num_shape * np.pi * (rp_max ** 2 - rp_min ** 2) * abs(pi_max - pi_min) / L3


def volume_spherical_cap(mur, r):
    """measure_IA_base.py:274-291."""
    return np.pi / 3.0 * r ** 3 * (2 + mur) * (1 - mur) ** 2


def random_pairs_rmu(r_max, r_min, mur_max, mur_min, L3, Num_position, Num_shape):
    """'cross' RR of an (r, mu) bin, with the (N_p - 1) factor (measure_IA_base.py:293-351)."""
    RR = ((Num_position - 1.0) * Num_shape
          * (volume_spherical_cap(mur_min, r_max) - volume_spherical_cap(mur_max, r_max)
             - (volume_spherical_cap(mur_min, r_min) - volume_spherical_cap(mur_max, r_min)))
          / L3)
    return abs(RR)


def rr_grid(mode, r_bins, second_bins, L3, Num_position, Num_shape):
    """The RR loops of measure_w_box_jk.py:220-227 / measure_m_box_jk.py:232-239."""
    nr, n2 = len(r_bins) - 1, len(second_bins) - 1
    RR = np.zeros((nr, n2))
    f = random_pairs_rppi if mode == "rppi" else random_pairs_rmu
    for i in np.arange(0, nr):
        for p in np.arange(0, n2):
            RR[i, p] = f(r_bins[i + 1], r_bins[i], second_bins[p + 1], second_bins[p], L3,
                         Num_position, Num_shape)
    return RR


def estimator(mode, counts, r_bins, second_bins, boxsize, Num_position, Num_shape, R,
              jk=None):
    """Assemble every dataset a JK counter writes (measure_w_box_jk.py:213-287).

    ``counts`` holds the brute sums (DD, SplusD, ScrossD, variance, DD_jk,
    SplusD_jk).  ``jk`` = dict(jk_pos, jk_shape, num_box, R_jk) or None.
    Returns a dict of named arrays; realisation i is under key ('jk', i).
    """
    L3 = boxsize ** 3
    RR = rr_grid(mode, r_bins, second_bins, L3, Num_position, Num_shape)
    out = {}
    out["xi_g_plus"] = counts["SplusD"] / RR
    out["xi_g_cross"] = counts["ScrossD"] / RR
    out["sigmasq"] = counts["variance"] / RR ** 2
    out["xi_gg"] = (counts["DD"] / RR) - 1
    out["RR"] = RR
    dsep = (r_bins[1:] - r_bins[:-1]) / 2.0
    out["r_centres"] = r_bins[:-1] + abs(dsep)
    d2 = (second_bins[1:] - second_bins[:-1]) / 2.0
    out["second_centres"] = second_bins[:-1] + abs(d2)
    if jk is not None:
        num_box = jk["num_box"]
        volume_jk = L3 * (num_box - 1) / num_box
        RR_jk = np.zeros((num_box,) + RR.shape)
        for b in np.arange(num_box):
            npj = len(np.where(jk["jk_pos"] != b)[0])
            nsj = len(np.where(jk["jk_shape"] != b)[0])
            RR_jk[b] = rr_grid(mode, r_bins, second_bins, volume_jk, npj, nsj)
        out["RR_jk"] = RR_jk
        R_jk = jk["R_jk"]
        SP, DD = counts["SplusD"], counts["DD"]
        for i in range(num_box):
            num = SP * (2 * R) - counts["SplusD_jk"][i]
            out[("jk", i)] = dict(
                xi_g_plus=num / (RR_jk[i] * 2 * R_jk[i]),
                SplusD=num / (2 * R_jk[i]),
                xi_gg=((DD - counts["DD_jk"][i]) / RR_jk[i]) - 1,
                DD=DD - counts["DD_jk"][i],
                RR=RR_jk[i])
    return out


def project_w(xi, pi_bins):
    """w = sum_pi xi |dpi| (measure_IA_base.py:553-562)."""
    dpi = (pi_bins[1:] - pi_bins[:-1])
    dpi = np.array([dpi] * len(xi[:, 0]))
    return np.sum(xi * abs(dpi), axis=1)


def multipole(xi, mu_centres, mu_r_bins, corr):
    """(2l+1)/2 (l-m)!/(l+m)! P_l^m(mu) xi dmu summed over mu (measure_IA_base.py:628-655).

    corr 'g_plus' uses l = m = 2, 'gg' uses l = m = 0, with scipy's lpmn.
    """
    from scipy.special import lpmn
    sab = 2 if corr == "g_plus" else 0
    l = sab
    nr = xi.shape[0]
    mu = np.array(list(mu_centres) * nr).reshape((nr, len(mu_centres)))
    L = np.zeros((nr, len(mu_centres)))
    for n in np.arange(nr):
        for m in np.arange(len(mu_centres)):
            L_m, _ = lpmn(l, sab, mu[n, m])
            L[n, m] = L_m[-1, -1]
    dmur = (mu_r_bins[1:] - mu_r_bins[:-1])
    dmu = np.array(list(dmur) * nr).reshape((nr, len(dmur)))
    m = ((2 * l + 1) / 2.0 * math.factorial(l - sab) / math.factorial(l + sab) * L * xi * dmu)
    return np.sum(m, axis=1)


def combine_jackknife(realisations, num_bins_r):
    """mean, std, cov with the (n-1)/n rule (measure_IA_base.py:478-498)."""
    num_box = len(realisations)
    mean = np.zeros(num_bins_r)
    for b in np.arange(0, num_box):
        mean += realisations[b]
    mean /= num_box
    cov = np.zeros((num_bins_r, num_bins_r))
    std = np.zeros(num_bins_r)
    for b in np.arange(0, num_box):
        std += (realisations[b] - mean) ** 2
        for i in np.arange(num_bins_r):
            cov[:, i] += (realisations[b] - mean) * (realisations[b][i] - mean[i])
    std *= (num_box - 1) / num_box
    std = np.sqrt(std)
    cov *= (num_box - 1) / num_box
    return mean, std, cov
