"""Pins of the lightcone oracle (oracle/lightcone.py) and of the host helpers
of the lightcone front-end.  CPU only.

The reference holds no lightcone fixtures or known-answer tests, so the
oracle is pinned by what the geometry and its definitions fix:
  * hand-placed pairs at a known (r_p, pi) with the shape axis along,
    across and diagonal to the separation: e+ = -e, +e, 0 and ex = 0, 0, -e
    (the reference's sign: get_ellipticity(-e, phi_axis - phi_sep),
    measure_w_lightcone.py:163-168);
  * an independent per-pair loop with Python floats and math.*;
  * leave-one-patch-out recounts by masking (the reference's definition,
    measure_jackknife.py:84-130) equal total - tally;
  * weights x 1/2 -> DD / 4; over_h scales r_p by h;
  * the rp_cut threshold against sqrt on neighbouring doubles;
  * the ccl-free distance against the Einstein-de Sitter closed form.
"""
import math

import numpy as np
import pytest

from measure_ia_amd import binning
from measure_ia_amd.lightcone import FlatLCDM, rotated_ellipticity, rp_cut_threshold
from oracle import lightcone as olc

R_BINS = np.logspace(np.log10(0.5), np.log10(15.0), 9)
PI_BINS = np.linspace(-20, 20, 11)


def _one_pair(dra_deg, ddec_deg, dchi, e1, e2, chi=1000.0, dec=0.0, mode="rppi", second=PI_BINS):
    return olc.count(mode, np.array([20.0]), np.array([dec]), np.array([chi]), np.ones(1),
                     np.array([20.0 + dra_deg]), np.array([dec + ddec_deg]), np.array([chi + dchi]),
                     np.ones(1), R_BINS, second, 0.5, 15.0, e1=np.array([e1]), e2=np.array([e2]))


def _deg(length, chi=1000.0):
    return length / (chi * math.pi / 180.0)


@pytest.mark.parametrize("dra, ddec, e1, e2, ep, ex", [
    (_deg(3.0), 0.0, 0.3, 0.0, -0.3, 0.0),    # axis along the separation
    (0.0, _deg(3.0), 0.3, 0.0, +0.3, 0.0),    # axis across the separation
    (_deg(3.0), 0.0, 0.0, 0.3, 0.0, -0.3),    # axis at 45 degrees
    (_deg(3.0), 0.0, 0.0, -0.3, 0.0, +0.3),   # axis at -45 degrees
])
def test_known_answers(dra, ddec, e1, e2, ep, ex):
    out = _one_pair(dra, ddec, 5.0, e1, e2)
    assert out["NP"][0].sum() == 1
    # r_p = 3 lands in the bin of log10(3), pi = 5 in the pi bin of 5
    ir = int(np.searchsorted(R_BINS, 3.0, side="right") - 1)
    ip = int(np.searchsorted(PI_BINS, 5.0, side="right") - 1)
    assert out["NP"][0, ir, ip] == 1
    assert out["SplusD"][0, ir, ip] == pytest.approx(ep, abs=1e-12)
    assert out["ScrossD"][0, ir, ip] == pytest.approx(ex, abs=1e-12)


def test_dx_uses_the_position_declination():
    # dx = dRA * chi_n * cos(dec_n): the same RA offset is a shorter r_p at dec 60
    dra = _deg(3.0)
    at0 = _one_pair(dra, 0.0, 5.0, 0.3, 0.0, dec=0.0)
    at60 = _one_pair(dra, 0.0, 5.0, 0.3, 0.0, dec=60.0)
    i0 = np.argwhere(at0["NP"][0])[0][0]
    i60 = np.argwhere(at60["NP"][0])[0][0]
    assert R_BINS[i60] <= 1.5 < R_BINS[i60 + 1] and R_BINS[i0] <= 3.0 < R_BINS[i0 + 1]


def test_out_of_domain_pairs():
    assert _one_pair(_deg(20.0), 0.0, 5.0, 0.3, 0.0)["NP"].sum() == 0    # r_p >= r_max
    assert _one_pair(_deg(3.0), 0.0, 25.0, 0.3, 0.0)["NP"].sum() == 0    # pi >= pi_max
    assert _one_pair(_deg(0.2), 0.0, 5.0, 0.3, 0.0)["NP"].sum() == 0     # r_p < r_min


def _scalar(ra, dec, chi, w, ras, decs, chis, ws, e1, e2, mode, second, rp_cut=0.0):
    """Per-pair restatement with Python floats (shares no code with the oracle)."""
    nr, n2 = len(R_BINS) - 1, len(second) - 1
    step = (math.log10(15.0) - math.log10(0.5)) / nr
    out = {}
    for n in range(len(ra)):
        cn = math.cos(float(dec[n]) / 180 * math.pi)
        for s in range(len(ras)):
            los = float(chis[s]) - float(chi[n])
            dx = (float(ras[s]) - float(ra[n])) / 180 * math.pi * float(chi[n]) * cn
            dy = (float(decs[s]) - float(dec[n])) / 180 * math.pi * float(chi[n])
            rp = math.sqrt(dx * dx + dy * dy)
            if mode == "rppi":
                r = rp
                if not (R_BINS[0] <= rp < R_BINS[-1] and second[0] <= los < second[-1]):
                    continue
                i2 = math.floor(los / ((second[-1] - second[0]) / n2) - second[0] / ((second[-1] - second[0]) / n2))
            else:
                r = math.sqrt(dx * dx + dy * dy + los * los)
                if not (rp > rp_cut and R_BINS[0] <= r < R_BINS[-1]):
                    continue
                i2 = math.floor((los / r) / (2.0 / n2) - second[0] / (2.0 / n2))
            ir = min(math.floor(math.log10(r) / step - math.log10(R_BINS[0]) / step), nr - 1)
            i2 = min(i2, n2 - 1)
            e = math.hypot(float(e1[s]), float(e2[s]))
            pa = 0.5 * math.atan2(float(e2[s]), float(e1[s]))
            ps = math.atan2(dy, dx)
            wp = float(w[n]) * float(ws[s])
            acc = out.setdefault((ir, i2), [0.0, 0.0, 0.0])
            acc[0] += wp
            acc[1] += wp * -e * math.cos(2 * (pa - ps))
            acc[2] += wp * -e * math.sin(2 * (pa - ps))
    return out


@pytest.mark.parametrize("mode", ["rppi", "rmu"])
def test_against_scalar_loop(mode):
    rng = np.random.default_rng(7)
    n, m = 60, 50
    ra, dec = rng.uniform(10, 11, n), rng.uniform(-0.5, 0.5, n)
    chi = rng.uniform(500, 520, n)
    ras, decs, chis = rng.uniform(10, 11, m), rng.uniform(-0.5, 0.5, m), rng.uniform(500, 520, m)
    w, ws = rng.uniform(0.5, 1.5, n), rng.uniform(0.5, 1.5, m)
    e1, e2 = rng.normal(0, 0.2, m), rng.normal(0, 0.2, m)
    second = PI_BINS if mode == "rppi" else np.linspace(-1, 1, 11)
    out = olc.count(mode, ra, dec, chi, w, ras, decs, chis, ws, R_BINS, second, 0.5, 15.0, e1=e1, e2=e2,
                    clamp=True)
    ref = _scalar(ra, dec, chi, w, ras, decs, chis, ws, e1, e2, mode, second)
    assert sum(v[0] > 0 for v in ref.values()) > 5
    for (ir, i2), (dd, sp, sx) in ref.items():
        assert out["DD"][0, ir, i2] == pytest.approx(dd, rel=1e-12)
        assert out["SplusD"][0, ir, i2] == pytest.approx(sp, rel=1e-9, abs=1e-12)
        assert out["ScrossD"][0, ir, i2] == pytest.approx(sx, rel=1e-9, abs=1e-12)
    assert out["DD"][0].sum() == pytest.approx(sum(v[0] for v in ref.values()), rel=1e-12)


def _sample(seed, n=400, m=300):
    rng = np.random.default_rng(seed)
    ra, dec, chi = rng.uniform(10, 14, n), rng.uniform(-2, 2, n), rng.uniform(300, 340, n)
    ras, decs, chis = rng.uniform(10, 14, m), rng.uniform(-2, 2, m), rng.uniform(300, 340, m)
    w, ws = rng.uniform(0.5, 1.5, n), rng.uniform(0.5, 1.5, m)
    e1, e2 = rng.normal(0, 0.2, m), rng.normal(0, 0.2, m)
    return (ra, dec, chi, w, ras, decs, chis, ws), e1, e2


def test_tallies_equal_masked_recounts():
    arrays, e1, e2 = _sample(3)
    ra, ras = arrays[0], arrays[4]
    P = 4
    patch = np.minimum(((ra - 10) / 4 * P).astype(int), P - 1)
    patch_s = np.minimum(((ras - 10) / 4 * P).astype(int), P - 1)
    full = olc.count("rppi", *arrays, R_BINS, PI_BINS, 0.5, 15.0, e1=e1, e2=e2, patch=patch,
                     patch_s=patch_s, n_patches=P)
    reals = olc.realisations("rppi", arrays, patch, patch_s, range(P), R_BINS, PI_BINS, 0.5, 15.0,
                             e1=e1, e2=e2)
    for i in range(P):
        np.testing.assert_array_equal(reals[i]["NP"][0], full["NP"][0] - full["NP"][1 + i])
        np.testing.assert_allclose(reals[i]["DD"][0], full["DD"][0] - full["DD"][1 + i], rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(reals[i]["SplusD"][0], full["SplusD"][0] - full["SplusD"][1 + i],
                                   rtol=1e-9, atol=1e-9)


def test_weight_scaling_and_over_h():
    arrays, e1, e2 = _sample(4)
    a = olc.count("rppi", *arrays, R_BINS, PI_BINS, 0.5, 15.0, e1=e1, e2=e2)
    half = list(arrays)
    half[3] = arrays[3] * 0.5
    half[7] = arrays[7] * 0.5
    b = olc.count("rppi", *half, R_BINS, PI_BINS, 0.5, 15.0, e1=e1, e2=e2)
    np.testing.assert_allclose(b["DD"], a["DD"] / 4, rtol=1e-14)
    np.testing.assert_allclose(b["SplusD"], a["SplusD"] / 4, rtol=1e-12, atol=1e-14)
    # over_h with h = 1 is the identity; with h = 2 every r_p doubles (pi unchanged)
    c = olc.count("rppi", *arrays, R_BINS, PI_BINS, 0.5, 15.0, e1=e1, e2=e2, h=1.0, over_h=True)
    np.testing.assert_array_equal(c["NP"], a["NP"])
    lo = olc.count("rppi", *arrays, R_BINS / 2, PI_BINS, 0.25, 7.5)
    hi = olc.count("rppi", *arrays, R_BINS, PI_BINS, 0.5, 15.0, h=2.0, over_h=True)
    assert np.abs(lo["NP"] - hi["NP"]).sum() <= 2  # equal up to ties from rounding at the edges


def test_rotated_ellipticity_closed_form():
    rng = np.random.default_rng(5)
    e1, e2 = rng.normal(0, 0.3, 1000), rng.normal(0, 0.3, 1000)
    c2, s2 = rotated_ellipticity(e1, e2)
    np.testing.assert_allclose(c2, -e1, atol=1e-15)
    np.testing.assert_allclose(s2, -e2, atol=1e-15)


@pytest.mark.parametrize("cut", [0.0, 0.5, 1.0, 2.3456789, 10.0])
def test_rp_cut_threshold(cut):
    t = rp_cut_threshold(cut)
    keys = binning.to_ordered(np.array([t])) + np.arange(-50, 51)
    s = binning.from_ordered(keys)
    s = s[s >= 0]
    np.testing.assert_array_equal(np.sqrt(s) > cut, s > t)
    assert rp_cut_threshold(None) == 0.0 and rp_cut_threshold(-1.0) < 0


def test_distance_einstein_de_sitter():
    # Omega_m = 1, no radiation: chi(z) = 2 c/H0 (1 - 1/sqrt(1+z))
    cosmo = FlatLCDM(Omega_c=0.955, Omega_b=0.045, h=0.7, T_CMB=0.0)
    z = np.array([0.01, 0.1, 0.5, 1.0, 3.0])
    ref = 2 * 299792.458 / 70.0 * (1 - 1 / np.sqrt(1 + z))
    np.testing.assert_allclose(cosmo.comoving_radial_distance(1 / (1 + z)), ref, rtol=1e-8)


def test_default_distance_low_z_and_radiation():
    cosmo = FlatLCDM()
    z = 1e-3
    # chi -> c z / H0 at low z
    assert cosmo.comoving_radial_distance(1 / (1 + z)) == pytest.approx(299792.458 / 70.0 * z, rel=2e-3)
    assert 1e-5 < cosmo.Or < 1e-4 and abs(cosmo.Om + cosmo.Or + cosmo.OL - 1) < 1e-15
