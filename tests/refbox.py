"""Expected HDF5 content of a reference measurement, built from the oracle.

Mirrors what MeasureIABox.measure_xi_w / measure_xi_multipoles write
(src/measureia/measure_IA.py:68-262): the counter datasets of the chosen
path (brute measure_w_box_jk.py:248-287, tree :498-537, multiprocessing
:816-853; no-jackknife measure_w_box.py:208-227; multipoles
measure_m_box_jk.py:261-300, measure_m_box.py:408-427), then w or the
multipoles of the main dataset and of every realisation, then the jackknife
mean/std/cov.  Returns {hdf5 path: array}.
"""
import numpy as np

from oracle import brute, post, prep


def expected_output(data, boxsize, snap_group, r_lim, nr, n2, pi_max, grp, ds, corr_type, num_jk,
                    layout, masks=None, ellipticity="distortion", periodic=True):
    pr = prep.prepare(data, masks, ellipticity)
    r_bins, pi_bins, mu_bins = prep.bin_edges(r_lim[0], r_lim[1], nr, n2, pi_max, boxsize)
    mode = "rppi" if grp == "w" else "rmu"
    sb = pi_bins if mode == "rppi" else mu_bins
    L_sub = int(round(num_jk ** (1 / 3))) if num_jk else 0
    jk = None
    jp = js = None
    if num_jk:
        jp, js = prep.jackknife_region_indices(pr["positions"], pr["shapes"], boxsize, L_sub)
    c = brute.count_pairs(pr, r_bins, sb, r_lim[0], r_lim[1], boxsize, mode=mode, jk_pos=jp,
                          jk_shape=js, num_box=num_jk, periodic=periodic)
    if num_jk:
        jk = dict(jk_pos=jp, jk_shape=js, num_box=num_jk,
                  R_jk=prep.responsivity_jk(pr["weight_shape"], pr["e"], js, num_jk))
    Np, Ns = len(pr["positions"]), len(pr["shapes"])
    est = post.estimator(mode, c, r_bins, sb, boxsize, Np, Ns, pr["R"], jk)
    out = {}
    sn = snap_group
    gname = "w" if grp == "w" else "multipoles"
    sec = "pi" if grp == "w" else "mu_r"
    rn = "rp" if grp == "w" else "r"
    rc, scn = est["r_centres"], est["second_centres"]

    def put(group, name, v):
        path = "/".join(p for p in f"{group}/{name}".split("/") if p)
        out[path] = np.asarray(v)

    jkg = f"{ds}_jk{num_jk}" if num_jk else ""
    if num_jk:
        sig = {"brute": est["sigmasq"], "tree": np.zeros_like(est["RR"]), "multiprocessing": None}[layout]
        g = f"{sn}/{gname}/xi_g_plus/"
        put(g, ds, est["xi_g_plus"]); put(g, ds + "_SplusD", c["SplusD"]); put(g, ds + "_RR_g_plus", est["RR"])
        if sig is not None:
            put(g, ds + "_sigmasq", sig)
        put(g, ds + f"_{rn}", rc); put(g, ds + f"_{sec}", scn)
        for i in range(num_jk):
            g = f"{sn}/{gname}/xi_g_plus/{jkg}"
            r = est[("jk", i)]
            put(g, ds + f"_{i}", r["xi_g_plus"]); put(g, ds + f"_{i}_SplusD", r["SplusD"])
            put(g, ds + f"_{i}_RR", r["RR"]); put(g, ds + f"_{i}_{rn}", rc); put(g, ds + f"_{i}_{sec}", scn)
        g = f"{sn}/{gname}/xi_g_cross/{jkg}"
        put(g, ds + "_ScrossD", c["ScrossD"]); put(g, ds, est["xi_g_cross"]); put(g, ds + "_RR_g_cross", est["RR"])
        if sig is not None:
            put(g, ds + "_sigmasq", sig)
        put(g, ds + f"_{rn}", rc); put(g, ds + f"_{sec}", scn)
        g = f"{sn}/{gname}/xi_gg/"
        put(g, ds, est["xi_gg"]); put(g, ds + "_DD", c["DD"]); put(g, ds + "_RR_gg", est["RR"])
        if sig is not None:
            put(g, ds + "_sigmasq", sig)
        put(g, ds + f"_{rn}", rc); put(g, ds + f"_{sec}", scn)
        for i in range(num_jk):
            g = f"{sn}/{gname}/xi_gg/{jkg}"
            r = est[("jk", i)]
            put(g, ds + f"_{i}", r["xi_gg"]); put(g, ds + f"_{i}_DD", r["DD"]); put(g, ds + f"_{i}_RR", r["RR"])
            put(g, ds + f"_{i}_{rn}", rc); put(g, ds + f"_{i}_{sec}", scn)
    else:
        for name, xi, raw, rr_name, raw_name in (("xi_g_plus", est["xi_g_plus"], c["SplusD"], "_RR_g_plus", "_SplusD"),
                                                 ("xi_g_cross", est["xi_g_cross"], c["ScrossD"], "_RR_g_cross", "_ScrossD"),
                                                 ("xi_gg", est["xi_gg"], c["DD"], "_RR_gg", "_DD")):
            g = f"{sn}/{gname}/{name}/"
            put(g, ds, xi); put(g, ds + raw_name, raw); put(g, ds + rr_name, est["RR"])
            put(g, ds + f"_{rn}", rc); put(g, ds + f"_{sec}", scn)
    # post-processing: the main dataset and every realisation
    names = [(ds, "")] + [(f"{ds}_{i}", jkg) for i in range(num_jk)]
    corrs = {"both": ["g_plus", "gg"], "g+": ["g_plus"], "gg": ["gg"]}[corr_type]
    for name, jg in names:
        for corr in corrs:
            xi = out["/".join(p for p in f"{sn}/{gname}/xi_{corr}/{jg}/{name}".split("/") if p)]
            if grp == "w":
                g = f"{sn}/w_{corr}/{jg}"
                put(g, name + "_rp", rc)
                put(g, name, post.project_w(xi, pi_bins))
            else:
                g = f"{sn}/multipoles_{corr}/{jg}"
                put(g, name + "_r", rc)
                put(g, name, post.multipole(xi, scn, mu_bins, corr))
    if num_jk:
        for corr in corrs:
            cg = f"w_{corr}" if grp == "w" else f"multipoles_{corr}"
            reals = [out["/".join(p for p in f"{sn}/{cg}/{jkg}/{ds}_{i}".split("/") if p)]
                     for i in range(num_jk)]
            mean, std, cov = post.combine_jackknife(reals, nr)
            put(f"{sn}/{cg}", f"{ds}_mean_{num_jk}", mean)
            put(f"{sn}/{cg}", f"{ds}_jackknife_{num_jk}", std)
            put(f"{sn}/{cg}", f"{ds}_jackknife_cov_{num_jk}", cov)
    return out
