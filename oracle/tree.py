"""KDTree + multiprocessing pair counter: the reference's fast CPU path (oracle).

Test infrastructure only: imported by tests/, smoke() and bench.py's
cpu_baseline, never by the product package.

Restates src/measureia/measure_w_box_jk.py:386-461 (single process) and
:543-646 + :761-780 (Pool over shape chunks), and the (r, mu) twin in
src/measureia/measure_m_box_jk.py:403-487: a periodic scipy KDTree over the
positions (projected coordinates for (r_p, pi), full 3-D for (r, mu)), per
chunk a KDTree over the shapes, query_ball_tree at r_min and r_max and their
set difference, then per shape the same numpy arithmetic as the brute
counter, with np.add.at scatter-adds and per-region tallies.

The temp-HDF5 staging of the reference's multiprocessing path (:716-730) is
not reproduced: workers inherit the arrays through fork, which makes this
baseline slightly faster than the reference.
"""
import multiprocessing as mp
import os
import sys
import time

import numpy as np
from scipy.spatial import KDTree

_G = {}


def _setup(prep, r_bins, second_bins, r_min, r_max, boxsize, mode, jk_pos, jk_shape, num_box,
           periodic, rp_cut):
    pos = prep["positions"]
    nl = prep["not_los"]
    tree_pts = pos[:, nl] if mode == "rppi" else pos
    _G.clear()
    _G.update(prep=prep, r_bins=r_bins, second_bins=second_bins, r_min=r_min, r_max=r_max,
              boxsize=boxsize, mode=mode, jk_pos=jk_pos, jk_shape=jk_shape, num_box=num_box,
              periodic=periodic, rp_cut=0.0 if rp_cut is None else rp_cut,
              pos_tree=KDTree(tree_pts, boxsize=boxsize))


def _chunk(bounds):
    """One shape chunk: measure_w_box_jk.py:543-646 (rppi) / measure_m_box_jk.py:570-682 (rmu)."""
    i, i2 = bounds
    g = _G
    prep = g["prep"]
    pos, shp, axis, e = prep["positions"], prep["shapes"], prep["axis"], prep["e"]
    w, ws, R = prep["weight"], prep["weight_shape"], prep["R"]
    los, nl = prep["los"], prep["not_los"]
    r_bins, sb = g["r_bins"], g["second_bins"]
    L, mode, num_box = g["boxsize"], g["mode"], g["num_box"]
    L_0p5 = L / 2.0
    nr, n2 = len(r_bins) - 1, len(sb) - 1
    log_step = (np.log10(g["r_max"]) - np.log10(g["r_min"])) / nr
    step2 = (sb[-1] - sb[0]) / n2 if mode == "rppi" else 2.0 / n2
    DD = np.zeros((nr, n2))
    SP = np.zeros((nr, n2))
    SX = np.zeros((nr, n2))
    DD_jk = np.zeros((max(num_box, 1), nr, n2))
    SP_jk = np.zeros((max(num_box, 1), nr, n2))
    absum = g["abs_sums"]
    SPA = np.zeros((nr, n2)) if absum else None
    SXA = np.zeros((nr, n2)) if absum else None
    SPA_jk = np.zeros((max(num_box, 1), nr, n2)) if absum else None
    npairs = 0
    shp_i = shp[i:i2]
    tree_pts = shp_i[:, nl] if mode == "rppi" else shp_i
    shape_tree = KDTree(tree_pts, boxsize=L)
    ind_min = shape_tree.query_ball_tree(g["pos_tree"], g["r_min"])
    ind_max = shape_tree.query_ball_tree(g["pos_tree"], g["r_max"])
    jk_pos, jk_shape = g["jk_pos"], g["jk_shape"]
    for n in range(len(shp_i)):
        idx = np.setdiff1d(ind_max[n], ind_min[n])
        if len(idx) == 0:
            continue
        s = i + n
        sep = shp[s] - pos[idx]
        if g["periodic"]:
            sep[sep > L_0p5] -= L
            sep[sep < -L_0p5] += L
        proj = sep[:, nl]
        lsep = sep[:, los]
        rp = np.sqrt(np.sum(proj ** 2, axis=1))
        with np.errstate(invalid="ignore", divide="ignore"):
            sdir = (proj.transpose() / rp).transpose()
            phi = np.arccos(sdir[:, 0] * axis[s, 0] + sdir[:, 1] * axis[s, 1])
        e_plus, e_cross = e[s] * np.cos(2 * phi), e[s] * np.sin(2 * phi)
        e_plus[np.isnan(e_plus)] = 0.0
        e_cross[np.isnan(e_cross)] = 0.0
        if mode == "rppi":
            sl = rp
            mask = (sl >= r_bins[0]) * (sl < r_bins[-1]) * (lsep >= sb[0]) * (lsep < sb[-1])
            x2 = lsep[mask]
        else:
            sl = np.sqrt(np.sum(sep ** 2, axis=1))
            with np.errstate(invalid="ignore", divide="ignore"):
                mu = lsep / sl
            mask = (rp > g["rp_cut"]) * (sl >= r_bins[0]) * (sl < r_bins[-1])
            x2 = mu[mask]
        ind_r = np.array(np.floor(np.log10(sl[mask]) / log_step - np.log10(r_bins[0]) / log_step), dtype=int)
        ind_2 = np.array(np.floor(x2 / step2 - sb[0] / step2), dtype=int)
        if np.any(ind_2 == n2):
            ind_2[ind_2 >= n2] -= 1
        if np.any(ind_r == nr):
            ind_r[ind_r >= nr] -= 1
        wp = w[idx][mask] * ws[s]
        np.add.at(SP, (ind_r, ind_2), (wp * e_plus[mask]) / (2 * R))
        np.add.at(SX, (ind_r, ind_2), (wp * e_cross[mask]) / (2 * R))
        np.add.at(DD, (ind_r, ind_2), wp)
        npairs += int(mask.sum())
        if absum:  # tolerance scales for the tests, not part of the timed baseline
            np.add.at(SPA, (ind_r, ind_2), np.abs((wp * e_plus[mask]) / (2 * R)))
            np.add.at(SXA, (ind_r, ind_2), np.abs((wp * e_cross[mask]) / (2 * R)))
        if num_box > 0:
            reg_p = jk_pos[idx][mask]
            pm = np.where(reg_p != jk_shape[s])[0]
            np.add.at(SP_jk, (jk_shape[s], ind_r, ind_2), wp * e_plus[mask])
            np.add.at(SP_jk, (reg_p[pm], ind_r[pm], ind_2[pm]), wp[pm] * e_plus[mask][pm])
            np.add.at(DD_jk, (jk_shape[s], ind_r, ind_2), wp)
            np.add.at(DD_jk, (reg_p[pm], ind_r[pm], ind_2[pm]), wp[pm])
            if absum:
                np.add.at(SPA_jk, (jk_shape[s], ind_r, ind_2), np.abs(wp * e_plus[mask]))
                np.add.at(SPA_jk, (reg_p[pm], ind_r[pm], ind_2[pm]), np.abs(wp[pm] * e_plus[mask][pm]))
    return SP, SX, DD, DD_jk, SP_jk, npairs, SPA, SXA, SPA_jk


def count_pairs(prep, r_bins, second_bins, r_min, r_max, boxsize, *, mode="rppi",
                jk_pos=None, jk_shape=None, num_box=0, periodic=True, rp_cut=None,
                chunk_size=1000, num_nodes=1, chunk_subset=None, abs_sums=False, progress=False):
    """Tree counter over all shape chunks, or over ``chunk_subset`` (chunk numbers).

    Returns the same keys as oracle.brute.count_pairs (without 'variance',
    which the reference's tree path never accumulates, and without the
    per-bin sums of |terms| unless ``abs_sums``), plus 'seconds' of wall time
    for the counting itself and 'chunks' = (done, total).
    """
    _setup(prep, r_bins, second_bins, r_min, r_max, boxsize, mode, jk_pos, jk_shape, num_box,
           periodic, rp_cut)
    _G["abs_sums"] = abs_sums
    Ns = len(prep["shapes"])
    starts = np.arange(0, Ns, chunk_size)
    bounds = [(int(s), int(min(Ns, s + chunk_size))) for s in starts]
    total = len(bounds)
    if chunk_subset is not None:
        bounds = [bounds[c] for c in chunk_subset]
    t0 = time.perf_counter()
    if num_nodes > 1 and len(bounds) > 1:
        ctx = mp.get_context("fork")
        with ctx.Pool(num_nodes) as p:
            if progress:
                # chunk results in order, a line about every tenth of the chunks
                result = []
                step = max(1, len(bounds) // 10)
                for r in p.imap(_chunk, bounds):
                    result.append(r)
                    if len(result) % step == 0:
                        print(f"tree oracle: {len(result)}/{len(bounds)} chunks, "
                              f"{time.perf_counter() - t0:.0f} s", file=sys.stderr, flush=True)
            else:
                result = p.map(_chunk, bounds)
    else:
        result = [_chunk(b) for b in bounds]
    secs = time.perf_counter() - t0
    nr, n2 = len(r_bins) - 1, len(second_bins) - 1
    SP = np.zeros((nr, n2))
    SX = np.zeros((nr, n2))
    DD = np.zeros((nr, n2))
    DD_jk = np.zeros((max(num_box, 1), nr, n2))
    SP_jk = np.zeros((max(num_box, 1), nr, n2))
    npairs = 0
    SPA, SXA, SPA_jk = np.zeros((nr, n2)), np.zeros((nr, n2)), np.zeros((max(num_box, 1), nr, n2))
    for r in result:
        SP += r[0]
        SX += r[1]
        DD += r[2]
        DD_jk += r[3]
        SP_jk += r[4]
        npairs += r[5]
        if abs_sums:
            SPA += r[6]
            SXA += r[7]
            SPA_jk += r[8]
    out = dict(DD=DD, SplusD=SP, ScrossD=SX, npairs=npairs, seconds=secs, chunks=(len(bounds), total))
    if abs_sums:
        out["SplusD_abs"], out["ScrossD_abs"] = SPA, SXA
    if num_box > 0:
        out["DD_jk"] = DD_jk
        out["SplusD_jk"] = SP_jk
        if abs_sums:
            out["SplusD_jk_abs"] = SPA_jk
    return out


def default_workers():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1
