"""Seeded lightcone counting cases shared by the oracle and GPU tests."""
import zlib

import numpy as np

from measure_ia_amd import binning, synth
from measure_ia_amd.engine import LightconeArrays
from measure_ia_amd.lightcone import rotated_ellipticity, rp_cut_threshold
from oracle import lightcone as olc


def _patches(ra, n, lo, hi):
    return np.clip(((ra - lo) / (hi - lo) * n).astype(int), 0, n - 1)


def make_case(name):
    """Returns (arrays dict, cfg)."""
    seed = zlib.crc32(name.encode())
    cfg = dict(mode="rppi", r_min=0.5, r_max=15.0, nr=8, n2=10, pi_max=20.0, shapes=True, h=0.7,
               over_h=False, rp_cut=None, P=0, weights=True, clamp=True)
    kw = dict(ra=(10.0, 25.0), dec=(-5.0, 5.0), chi=(300.0, 900.0))
    n_pos, n_shape = 1500, 800
    if name == "rppi_shapes_patches":
        cfg.update(P=6)
    elif name == "rmu_shapes_patches":
        cfg.update(mode="rmu", P=5, clamp=False)
    elif name == "rppi_counts_overh":
        cfg.update(shapes=False, over_h=True, P=4)
    elif name == "rmu_counts_rpcut":
        cfg.update(mode="rmu", shapes=False, rp_cut=1.0, P=3, clamp=False)
    elif name == "rmu_shapes_overh_rpcut":
        cfg.update(mode="rmu", over_h=True, rp_cut=2.0, clamp=False)
    elif name == "high_dec_wide":
        kw = dict(ra=(0.0, 120.0), dec=(50.0, 80.0), chi=(40.0, 400.0))
        cfg.update(P=4, r_max=10.0)
    elif name == "auto_unweighted":
        cfg.update(weights=False, P=3)
    elif name == "nan_ellipticity":
        cfg.update(P=2)
    else:
        raise KeyError(name)
    a = synth.lightcone(n_pos, n_shape, seed=seed % (2 ** 31), weights=cfg["weights"], **kw)
    if name == "auto_unweighted":
        # shape sample = position sample: every object pairs with itself at r = 0
        for k in ("RA", "DEC", "chi"):
            a[f"{k}_shape_sample"] = a[k].copy()
        rng = np.random.default_rng(seed)
        e12 = synth.ellipticity_noise(rng, n_pos, 0.25)
        a["e1"], a["e2"] = e12[:, 0], e12[:, 1]
    if name == "nan_ellipticity":
        a["e1"][::17] = np.nan
    if not cfg["weights"]:
        a["weight"] = np.ones(len(a["RA"]))
        a["weight_shape_sample"] = np.ones(len(a["RA_shape_sample"]))
    if cfg["over_h"]:
        a["chi"] = a["chi"] * cfg["h"]
        a["chi_shape_sample"] = a["chi_shape_sample"] * cfg["h"]
    if cfg["P"]:
        lo, hi = kw["ra"]
        a["patch"] = _patches(a["RA"], cfg["P"], lo, hi)
        a["patch_s"] = _patches(a["RA_shape_sample"], cfg["P"], lo, hi)
    return a, cfg


CASES = ["rppi_shapes_patches", "rmu_shapes_patches", "rppi_counts_overh", "rmu_counts_rpcut",
         "rmu_shapes_overh_rpcut", "high_dec_wide", "auto_unweighted", "nan_ellipticity"]


def bins(cfg):
    r_bins = np.logspace(np.log10(cfg["r_min"]), np.log10(cfg["r_max"]), cfg["nr"] + 1)
    if cfg["mode"] == "rppi":
        second = np.linspace(-cfg["pi_max"], cfg["pi_max"], cfg["n2"] + 1)
        th = binning.rppi_thresholds(r_bins, second, cfg["r_min"], cfg["r_max"])
    else:
        second = np.linspace(-1, 1, cfg["n2"] + 1)
        th = binning.rmu_thresholds(r_bins, second, cfg["r_min"], cfg["r_max"])
    return r_bins, second, th


def device_arrays(a, cfg):
    shapes = cfg["shapes"]
    c2 = s2 = None
    if shapes:
        c2, s2 = rotated_ellipticity(a["e1"], a["e2"])
    lc = LightconeArrays(ra=a["RA"], dec=a["DEC"], chi=a["chi"], cos_dec=np.cos(a["DEC"] / 180 * np.pi),
                         ra_s=a["RA_shape_sample"], dec_s=a["DEC_shape_sample"], chi_s=a["chi_shape_sample"],
                         weight=a["weight"], weight_shape=a["weight_shape_sample"], c2=c2, s2=s2,
                         proj_scale=cfg["h"] if cfg["over_h"] else 1.0,
                         rp2_cut=rp_cut_threshold(cfg["rp_cut"]) if cfg["mode"] == "rmu" else 0.0)
    if cfg["P"]:
        lc.patch = a["patch"].astype(np.int32)
        lc.patch_s = a["patch_s"].astype(np.int32)
        lc.n_patches = cfg["P"]
    return lc


def oracle_counts(a, cfg):
    r_bins, second, _ = bins(cfg)
    return olc.count(cfg["mode"], a["RA"], a["DEC"], a["chi"], a["weight"], a["RA_shape_sample"],
                     a["DEC_shape_sample"], a["chi_shape_sample"], a["weight_shape_sample"], r_bins, second,
                     cfg["r_min"], cfg["r_max"], e1=a["e1"] if cfg["shapes"] else None,
                     e2=a["e2"] if cfg["shapes"] else None, h=cfg["h"], over_h=cfg["over_h"],
                     rp_cut=cfg["rp_cut"], clamp=cfg["clamp"], patch=a.get("patch"),
                     patch_s=a.get("patch_s"), n_patches=cfg["P"])
