"""Input preparation, bin edges and jackknife region ids (oracle).

Test infrastructure only: imported by tests/, smoke() and bench.py's
cpu_baseline, never by the product package.
"""
import numpy as np


def bin_edges(r_min, r_max, num_bins_r, num_bins_pi, pi_max, boxsize):
    """Bin edges as MeasureIABase.__init__ builds them.

    src/measureia/measure_IA_base.py:155-167: r_bins = logspace(log10 rmin,
    log10 rmax, nr+1); pi_max defaults to L/2; pi_bins = linspace(-pi_max,
    pi_max, npi+1); mu_r_bins = linspace(-1, 1, npi+1).
    """
    r_bins = np.logspace(np.log10(r_min), np.log10(r_max), num_bins_r + 1)
    if pi_max is None:
        if boxsize is None:
            raise ValueError("Both pi_max and boxsize are None.")
        pi_max = boxsize / 2.0
    pi_bins = np.linspace(-pi_max, pi_max, num_bins_pi + 1)
    mu_r_bins = np.linspace(-1, 1, num_bins_pi + 1)
    return r_bins, pi_bins, mu_r_bins


def prepare(data, masks=None, ellipticity="distortion"):
    """Selected arrays, e and R, as every box counter prepares them.

    Follows src/measureia/measure_w_box_jk.py:97-139 (identical blocks in the
    other counters): masks select rows; Axis_Direction is normalised; a
    missing weight mask selects the FIRST k weights (:113-122); e from q by
    the 'distortion' (1-q^2)/(1+q^2) or 'ellipticity' (1-q)/(1+q)
    definition; R = sum(w_s (1 - e^2/2)) / sum(w_s) with Python's sequential
    builtin sum.
    """
    n_pos = len(data["Position"])
    n_shape = len(data["Position_shape_sample"])
    weight_all = data.get("weight", np.ones(n_pos))
    weight_shape_all = data.get("weight_shape_sample", np.ones(n_shape))
    if masks is None:
        positions = data["Position"]
        shapes = data["Position_shape_sample"]
        axis_v = data["Axis_Direction"]
        q = data["q"]
        weight = weight_all
        weight_shape = weight_shape_all
    else:
        positions = data["Position"][masks["Position"]]
        shapes = data["Position_shape_sample"][masks["Position_shape_sample"]]
        axis_v = data["Axis_Direction"][masks["Axis_Direction"]]
        q = data["q"][masks["q"]]
        if "weight" in masks:
            wm = masks["weight"]
        else:
            wm = np.ones(n_pos, dtype=bool)
            wm[sum(masks["Position"]):n_pos] = 0
        if "weight_shape_sample" in masks:
            wsm = masks["weight_shape_sample"]
        else:
            wsm = np.ones(n_shape, dtype=bool)
            wsm[sum(masks["Position_shape_sample"]):n_shape] = 0
        weight = weight_all[wm]
        weight_shape = weight_shape_all[wsm]
    axis_len = np.sqrt(np.sum(axis_v ** 2, axis=1))
    axis = (axis_v.transpose() / axis_len).transpose()
    if ellipticity == "distortion":
        e = (1 - q ** 2) / (1 + q ** 2)
    elif ellipticity == "ellipticity":
        e = (1 - q) / (1 + q)
    else:
        raise ValueError("Invalid value for ellipticity. Choose 'distortion' or 'ellipticity'.")
    R = sum(weight_shape * (1 - e ** 2 / 2.0)) / sum(weight_shape)
    los = data["LOS"]
    not_los = np.array([0, 1, 2])[np.isin([0, 1, 2], los, invert=True)]
    return dict(positions=np.asarray(positions, dtype=np.float64),
                shapes=np.asarray(shapes, dtype=np.float64),
                axis=axis, e=e, weight=np.asarray(weight, dtype=np.float64),
                weight_shape=np.asarray(weight_shape, dtype=np.float64),
                R=R, los=los, not_los=not_los)


def jackknife_region_indices(positions, shapes, boxsize, L_subboxes):
    """Region id of every object, the reference's literal triple loop.

    src/measureia/measure_IA_base.py:404-452: L_sub = 2*L_0p5/L_subboxes;
    region i*L^2 + j*L + k for STRICT i*L_sub < x < (i+1)*L_sub on each axis;
    objects on a boundary (or outside) keep region 0.
    """
    L_0p5 = boxsize / 2.0
    L_sub = L_0p5 * 2.0 / L_subboxes
    jk_pos = np.zeros(len(positions))
    jk_shape = np.zeros(len(shapes))
    num_box = 0
    for i in np.arange(0, L_subboxes):
        for j in np.arange(0, L_subboxes):
            for k in np.arange(0, L_subboxes):
                xb = [i * L_sub, (i + 1) * L_sub]
                yb = [j * L_sub, (j + 1) * L_sub]
                zb = [k * L_sub, (k + 1) * L_sub]
                mp = ((positions[:, 0] > xb[0]) * (positions[:, 0] < xb[1])
                      * (positions[:, 1] > yb[0]) * (positions[:, 1] < yb[1])
                      * (positions[:, 2] > zb[0]) * (positions[:, 2] < zb[1]))
                ms = ((shapes[:, 0] > xb[0]) * (shapes[:, 0] < xb[1])
                      * (shapes[:, 1] > yb[0]) * (shapes[:, 1] < yb[1])
                      * (shapes[:, 2] > zb[0]) * (shapes[:, 2] < zb[1]))
                jk_pos[mp] = num_box
                jk_shape[ms] = num_box
                num_box += 1
    return np.array(jk_pos, dtype=int), np.array(jk_shape, dtype=int)


def jackknife_region_indices_by_axis(positions, shapes, boxsize, L_subboxes):
    """The same region ids with the triple loop factored per axis.

    The reference's box test (measure_IA_base.py:404-452) is a product of one
    strict interval test per axis, with the same bounds i*L_sub and
    (i+1)*L_sub for every box, so an object's box index on each axis can be
    found on its own: 3*L passes over the arrays instead of L^3.  An object
    that passes no interval on some axis keeps region 0, as in the loop.
    """
    L_0p5 = boxsize / 2.0
    L_sub = L_0p5 * 2.0 / L_subboxes

    def ids(x):
        idx = np.full(x.shape, -1, dtype=np.int64)
        for i in np.arange(0, L_subboxes):
            m = (x > i * L_sub) * (x < (i + 1) * L_sub)
            idx[m] = i
        ok = np.all(idx >= 0, axis=1)
        reg = idx[:, 0] * L_subboxes * L_subboxes + idx[:, 1] * L_subboxes + idx[:, 2]
        return np.where(ok, reg, 0).astype(int)

    return ids(np.asarray(positions)), ids(np.asarray(shapes))


def responsivity_jk(weight_shape, e, jk_shape, num_box):
    """R_jk[i] over shapes outside region i (measure_w_box_jk.py:213-216)."""
    R_jk = np.zeros(num_box)
    for i in np.arange(num_box):
        m = np.where(jk_shape != i)
        R_jk[i] = sum(weight_shape[m] * (1 - e[m] ** 2 / 2.0)) / sum(weight_shape[m])
    return R_jk
