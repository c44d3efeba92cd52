"""Seeded pair-counting cases shared by the oracle and GPU parity tests."""
import zlib

import numpy as np

from measure_ia_amd import binning, hostprep
from measure_ia_amd.engine import SampleArrays
from measure_ia_amd.synth import clustered, ellipticity_noise, shapes_from_components, wrap_box
from oracle import brute, prep


def make_case(name):
    """Return (data, boxsize, cfg) where cfg holds binning and jackknife settings."""
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    cfg = dict(r_min=0.1, r_max=10.0, nr=10, n2=8, pi_max=20.0, L_sub=3, los=2, weights=False,
               periodic=True)
    if name == "uniform_jk27":
        L = 60.0
        pos = wrap_box(rng.uniform(0, L, (3000, 3)), L)
        shp = wrap_box(rng.uniform(0, L, (1500, 3)), L)
    elif name == "uniform_auto_nojk":
        L = 50.0
        pos = wrap_box(rng.uniform(0, L, (2500, 3)), L)
        shp = pos
        cfg.update(L_sub=0)
    elif name == "clustered_jk8":
        L = 80.0
        cat = clustered(3000, 1200, L, 40, 1.0, 40, 0.2, 0.2, seed=11)
        pos = cat.data["Position"]
        shp = cat.data["Position_shape_sample"]
        cfg.update(L_sub=2, r_max=15.0, pi_max=30.0)
    elif name == "weighted_los0":
        L = 60.0
        pos = wrap_box(rng.uniform(0, L, (2500, 3)), L)
        shp = wrap_box(rng.uniform(0, L, (1200, 3)), L)
        cfg.update(weights=True, los=0)
    elif name == "los1_fullpi":
        L = 40.0
        pos = wrap_box(rng.uniform(0, L, (2000, 3)), L)
        shp = wrap_box(rng.uniform(0, L, (1000, 3)), L)
        cfg.update(los=1, pi_max=None, r_max=12.0)
    elif name == "lattice_boundaries":
        # integer lattice: exact ties with pi / mu edges and objects exactly on
        # jackknife faces (region 0 quirk) and at 0
        L = 12.0
        g = np.arange(0, 12, 1.0)
        X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
        pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
        shp = pos[rng.choice(len(pos), 600, replace=False)]
        cfg.update(r_min=1.0, r_max=5.0, nr=6, n2=10, pi_max=5.0, L_sub=3)
    elif name == "open_box_jk27":
        # periodicity=False: no minimum-image wrap, pairs across the box faces
        # fall outside the domain (measure_w_box_jk.py:160-162)
        L = 30.0
        pos = wrap_box(rng.uniform(0, L, (2500, 3)), L)
        shp = wrap_box(rng.uniform(0, L, (1200, 3)), L)
        cfg.update(periodic=False, r_max=12.0, pi_max=14.0)
    elif name == "reach_beyond_half_box":
        # r_max > L/2: every column and slab is reached once and separations
        # wrap on any axis (minimum image, measure_w_box_jk.py:160-162)
        L = 10.0
        pos = wrap_box(rng.uniform(0, L, (700, 3)), L)
        shp = wrap_box(rng.uniform(0, L, (300, 3)), L)
        cfg.update(r_min=0.3, r_max=7.0, nr=6, n2=6, pi_max=None, L_sub=2)
    elif name == "degenerate_axes":
        # zero-length axis vectors normalise to NaN; their pairs count in DD
        # with e+ = ex = 0 (the reference's isnan rule, measure_w_box_jk.py:174-175)
        L = 40.0
        pos = wrap_box(rng.uniform(0, L, (1500, 3)), L)
        shp = wrap_box(rng.uniform(0, L, (700, 3)), L)
        cfg.update(r_max=8.0, pi_max=10.0, L_sub=2)
    elif name == "tiny_box_wrap":
        L = 9.0
        pos = wrap_box(rng.uniform(0, L, (800, 3)), L)
        shp = wrap_box(rng.uniform(0, L, (500, 3)), L)
        cfg.update(r_min=0.5, r_max=4.4, nr=5, n2=6, pi_max=4.5, L_sub=2)
    elif name == "parallel_axes":
        # every shape's axis lies exactly along the projected separation to two
        # partner positions (one on each side), all on an integer lattice, so
        # d = (dx/r_p) a_x + (dy/r_p) a_y rounds to |d| > 1 for about a tenth of
        # those pairs and the reference's arccos gives NaN -> e+ = ex = 0
        # (measure_w_box_jk.py:166-175)
        L = 24.0
        shp = rng.integers(0, 24, (1500, 3)).astype(np.float64)
        sep = np.zeros((len(shp), 3))
        while True:
            bad = (sep[:, 0] == 0) & (sep[:, 1] == 0)
            if not bad.any():
                break
            sep[bad, :2] = rng.integers(-6, 7, (int(bad.sum()), 2))
        sep[:, 2] = rng.integers(-5, 6, len(shp))
        partners = np.concatenate([shp - sep, shp + sep])
        background = rng.integers(0, 24, (500, 3)).astype(np.float64)
        pos = wrap_box(np.concatenate([partners, background]), L)
        cfg.update(r_min=0.5, r_max=10.0, nr=8, n2=8, pi_max=8.0, L_sub=2)
    else:
        raise KeyError(name)
    e12 = ellipticity_noise(rng, len(shp), 0.25)
    axis, q = shapes_from_components(e12)
    axis = axis * rng.uniform(0.5, 2.0, (len(shp), 1))  # unnormalised on input
    if name == "degenerate_axes":
        axis[rng.choice(len(shp), len(shp) // 20, replace=False)] = 0.0
    if name == "parallel_axes":
        axis = sep[:, :2].copy()
    data = {"Position": pos, "Position_shape_sample": shp, "Axis_Direction": axis,
            "LOS": cfg["los"], "q": q}
    if cfg["weights"]:
        data["weight"] = rng.uniform(0.5, 1.5, len(pos))
        data["weight_shape_sample"] = rng.uniform(0.5, 1.5, len(shp))
    return data, L, cfg


CASES = ["uniform_jk27", "uniform_auto_nojk", "clustered_jk8", "weighted_los0", "los1_fullpi",
         "lattice_boundaries", "tiny_box_wrap", "open_box_jk27", "reach_beyond_half_box",
         "degenerate_axes", "parallel_axes"]


def oracle_counts(data, L, cfg, mode):
    pr = prep.prepare(data)
    r_bins, pi_bins, mu_bins = prep.bin_edges(cfg["r_min"], cfg["r_max"], cfg["nr"], cfg["n2"],
                                              cfg["pi_max"], L)
    sb = pi_bins if mode == "rppi" else mu_bins
    K = cfg["L_sub"] ** 3 if cfg["L_sub"] else 0
    jp = js = None
    if K:
        jp, js = prep.jackknife_region_indices(pr["positions"], pr["shapes"], L, cfg["L_sub"])
    out = brute.count_pairs(pr, r_bins, sb, cfg["r_min"], cfg["r_max"], L, mode=mode,
                            jk_pos=jp, jk_shape=js, num_box=K, periodic=cfg["periodic"])
    return pr, out, (r_bins, pi_bins, mu_bins)


def engine_sample(data, L, cfg):
    p = hostprep.prepare(data)
    n = cfg["L_sub"]
    return p, SampleArrays(positions=p.positions, shapes=p.shapes, axis=p.axis, e=p.e,
                           weight=p.weight if cfg["weights"] else None,
                           weight_shape=p.weight_shape if cfg["weights"] else None,
                           boxsize=L, los=p.los, jk_per_side=n,
                           jk_sub=(L / 2.0) * 2.0 / n if n else 0.0, periodic=cfg["periodic"])


def thresholds(cfg, L, mode):
    r_bins, pi_bins, mu_bins = prep.bin_edges(cfg["r_min"], cfg["r_max"], cfg["nr"], cfg["n2"],
                                              cfg["pi_max"], L)
    if mode == "rppi":
        return binning.rppi_thresholds(r_bins, pi_bins, cfg["r_min"], cfg["r_max"])
    return binning.rmu_thresholds(r_bins, mu_bins, cfg["r_min"], cfg["r_max"])


def rel_err(got, ref, scale):
    """|got - ref| / max(|ref|, sum|terms|), 0 where both vanish."""
    den = np.maximum(np.abs(ref), scale)
    d = np.abs(got - ref)
    return np.where(den > 0, d / np.where(den > 0, den, 1), d)


def compare(res, ref, R, K, weighted, tol=1e-6):
    """Check a device CountResult against oracle.brute output; returns max errors."""
    np.testing.assert_array_equal(res.npairs[0], ref["NP"])
    if weighted:
        e_dd = rel_err(res.dd[0], ref["DD"], np.abs(ref["DD"])).max()
        assert e_dd <= tol, e_dd
    else:
        np.testing.assert_array_equal(res.dd[0], ref["DD"])
    e_sp = rel_err(res.splus[0] / (2 * R), ref["SplusD"], ref["SplusD_abs"]).max()
    e_sx = rel_err(res.scross / (2 * R), ref["ScrossD"], ref["ScrossD_abs"]).max()
    assert e_sp <= tol, ("S+D", e_sp)
    assert e_sx <= tol, ("SxD", e_sx)
    errs = dict(splus=e_sp, scross=e_sx)
    if K:
        np.testing.assert_array_equal(res.npairs[1:], ref["NP_jk"])
        if weighted:
            # weighted DD tallies: sums of positive terms, plain relative error
            e_ddjk = rel_err(res.dd[1:], ref["DD_jk"], np.abs(ref["DD_jk"])).max()
            assert e_ddjk <= tol, ("DD_jk", e_ddjk)
            errs["dd_jk"] = e_ddjk
        else:
            np.testing.assert_array_equal(res.dd[1:], ref["DD_jk"])
        e_jk = rel_err(res.splus[1:], ref["SplusD_jk"], ref["SplusD_jk_abs"]).max()
        assert e_jk <= tol, ("S+D_jk", e_jk)
        errs["splus_jk"] = e_jk
    return errs


def dyadic_weights(rng, n):
    """Weights from {0.5, 1, 2}: every product w_p w_s and every partial sum of
    them is exact in FP64, so weighted DD must match the oracle bit for bit
    whatever the summation order."""
    return rng.choice(np.array([0.5, 1.0, 2.0]), n)
