"""Brute-force FP64 pair counters (oracle).

Test infrastructure only: imported by tests/, smoke() and bench.py's
cpu_baseline, never by the product package.

Every pair (shape s, position p) is evaluated with the reference's own
arithmetic, in its own op order, including np.log10 binning; the loop runs
over blocks of positions and is vectorised over all shapes, so the per-pair
operations are identical to the reference's loop over single positions.
Histograms are accumulated with np.bincount: DD with unit weights is exact
whatever the order; floating sums differ from np.add.at's order only in the
last bits.
"""
import numpy as np

_BLOCK_ELEMS = 1 << 21


def _wrap(sep, boxsize, periodic):
    """sep > L/2 -> sep - L, then sep < -L/2 -> sep + L (measure_w_box_jk.py:160-162)."""
    if periodic:
        L_0p5 = boxsize / 2.0
        sep[sep > L_0p5] -= boxsize
        sep[sep < -L_0p5] += boxsize
    return sep


def _ellipticity_components(dir0, dir1, axis, e):
    """e+ = e cos 2phi, ex = e sin 2phi with phi = arccos(d), NaN -> 0.

    measure_IA_base.py:187-227 and measure_w_box_jk.py:167-175: the dot
    product accumulates 0 + d0*a0 then + d1*a1.
    """
    with np.errstate(invalid="ignore"):
        dot = np.zeros(dir0.shape)
        dot += dir0 * axis[:, 0]
        dot += dir1 * axis[:, 1]
        phi = np.arccos(dot)
    e_plus = e * np.cos(2 * phi)
    e_cross = e * np.sin(2 * phi)
    e_plus[np.isnan(e_plus)] = 0.0
    e_cross[np.isnan(e_cross)] = 0.0
    return e_plus, e_cross


def _as_index(ind, n):
    """int cast + the reference's clamp (ind == n -> n-1) + numpy's negative wrap."""
    ind = np.array(ind, dtype=int)
    if np.any(ind == n):
        ind[ind >= n] -= 1
    ind = np.where(ind < 0, ind + n, ind)
    if ind.size and (ind.min() < 0 or ind.max() >= n):
        raise IndexError("bin index out of range")
    return ind


def count_pairs(prep, r_bins, second_bins, r_min, r_max, boxsize, *, mode="rppi",
                jk_pos=None, jk_shape=None, num_box=0, periodic=True, rp_cut=None):
    """Brute pair counter for (r_p, pi) ('rppi') or (r, mu_r) ('rmu') bins.

    rppi follows measure_w_box_jk.py:142-211 (no-JK: measure_w_box.py:136-184);
    rmu follows measure_m_box_jk.py:147-223 (no-JK: measure_m_box.py:338-383).
    Returns raw sums: DD, SplusD = sum w_p w_s e+/(2R), ScrossD, variance =
    sum (w_p w_s e+/(2R))^2, DD_jk and SplusD_jk (w_p w_s e+ without 1/(2R))
    per region as the reference tallies them, and npairs (integer count).
    """
    pos = prep["positions"]
    shp = prep["shapes"]
    axis = prep["axis"]
    e = prep["e"]
    w = prep["weight"]
    ws = prep["weight_shape"]
    R = prep["R"]
    los = prep["los"]
    nl = prep["not_los"]
    nr = len(r_bins) - 1
    n2 = len(second_bins) - 1
    log_step = (np.log10(r_max) - np.log10(r_min)) / nr
    if mode == "rppi":
        step2 = (second_bins[-1] - second_bins[0]) / n2
    elif mode == "rmu":
        step2 = 2.0 / n2
        if rp_cut is None:
            rp_cut = 0.0
    else:
        raise ValueError(mode)
    nb = nr * n2
    DD = np.zeros(nb)
    SP = np.zeros(nb)
    SX = np.zeros(nb)
    VAR = np.zeros(nb)
    SPA = np.zeros(nb)
    SXA = np.zeros(nb)
    NP = np.zeros(nb, dtype=np.int64)
    use_jk = num_box > 0
    DD_jk = np.zeros(num_box * nb)
    SP_jk = np.zeros(num_box * nb)
    SPA_jk = np.zeros(num_box * nb)
    NP_jk = np.zeros(num_box * nb, dtype=np.int64)
    npairs = 0
    Ns = len(shp)
    B = max(1, _BLOCK_ELEMS // max(Ns, 1))
    for b0 in range(0, len(pos), B):
        pb = pos[b0:b0 + B]
        sep = shp[None, :, :] - pb[:, None, :]
        _wrap(sep, boxsize, periodic)
        pa = sep[:, :, nl[0]]
        pbv = sep[:, :, nl[1]]
        lsep = sep[:, :, los]
        rp = np.sqrt(pa * pa + pbv * pbv)
        with np.errstate(invalid="ignore", divide="ignore"):
            d0 = pa / rp
            d1 = pbv / rp
        e_plus, e_cross = _ellipticity_components(d0, d1, axis, e[None, :])
        if mode == "rppi":
            sl = rp
            mask = (rp >= r_bins[0]) * (rp < r_bins[-1]) * (lsep >= second_bins[0]) * (lsep < second_bins[-1])
        else:
            sl = np.sqrt(sep[:, :, 0] * sep[:, :, 0] + sep[:, :, 1] * sep[:, :, 1] + sep[:, :, 2] * sep[:, :, 2])
            with np.errstate(invalid="ignore", divide="ignore"):
                mu = lsep / sl
            mask = (rp > rp_cut) * (sl >= r_bins[0]) * (sl < r_bins[-1])
        if not mask.any():
            continue
        ip, iS = np.nonzero(mask)
        ind_r = np.floor(np.log10(sl[mask]) / log_step - np.log10(r_bins[0]) / log_step)
        if mode == "rppi":
            ind_2 = np.floor(lsep[mask] / step2 - second_bins[0] / step2)
        else:
            ind_2 = np.floor(mu[mask] / step2 - second_bins[0] / step2)
        ind_r = _as_index(ind_r, nr)
        ind_2 = _as_index(ind_2, n2)
        flat = ind_r * n2 + ind_2
        wprod = w[b0 + ip] * ws[iS]
        ep = e_plus[mask]
        ex = e_cross[mask]
        DD += np.bincount(flat, weights=wprod, minlength=nb)
        SP += np.bincount(flat, weights=(wprod * ep) / (2 * R), minlength=nb)
        SX += np.bincount(flat, weights=(wprod * ex) / (2 * R), minlength=nb)
        VAR += np.bincount(flat, weights=((wprod * ep) / (2 * R)) ** 2, minlength=nb)
        SPA += np.bincount(flat, weights=np.abs((wprod * ep) / (2 * R)), minlength=nb)
        SXA += np.bincount(flat, weights=np.abs((wprod * ex) / (2 * R)), minlength=nb)
        npairs += len(flat)
        NP += np.bincount(flat, minlength=nb)
        if use_jk:
            rp_ = jk_pos[b0 + ip]
            rs_ = jk_shape[iS]
            raw = wprod * ep
            # tally[region of the position] for every pair ...
            DD_jk += np.bincount(rp_ * nb + flat, weights=wprod, minlength=num_box * nb)
            SP_jk += np.bincount(rp_ * nb + flat, weights=raw, minlength=num_box * nb)
            SPA_jk += np.bincount(rp_ * nb + flat, weights=np.abs(raw), minlength=num_box * nb)
            NP_jk += np.bincount(rp_ * nb + flat, minlength=num_box * nb)
            # ... and tally[region of the shape] when it differs.
            o = rs_ != rp_
            DD_jk += np.bincount((rs_ * nb + flat)[o], weights=wprod[o], minlength=num_box * nb)
            SP_jk += np.bincount((rs_ * nb + flat)[o], weights=raw[o], minlength=num_box * nb)
            SPA_jk += np.bincount((rs_ * nb + flat)[o], weights=np.abs(raw[o]), minlength=num_box * nb)
            NP_jk += np.bincount((rs_ * nb + flat)[o], minlength=num_box * nb)
    out = dict(DD=DD.reshape(nr, n2), SplusD=SP.reshape(nr, n2), ScrossD=SX.reshape(nr, n2),
               variance=VAR.reshape(nr, n2), npairs=npairs,
               SplusD_abs=SPA.reshape(nr, n2), ScrossD_abs=SXA.reshape(nr, n2),
               NP=NP.reshape(nr, n2))
    if use_jk:
        out["DD_jk"] = DD_jk.reshape(num_box, nr, n2)
        out["SplusD_jk"] = SP_jk.reshape(num_box, nr, n2)
        out["SplusD_jk_abs"] = SPA_jk.reshape(num_box, nr, n2)
        out["NP_jk"] = NP_jk.reshape(num_box, nr, n2)
    return out


def count_pairs_scalar(prep, r_bins, second_bins, r_min, r_max, boxsize, *, mode="rppi",
                       periodic=True, rp_cut=0.0):
    """Pure-Python per-pair loop (tiny inputs only).

    An independent restatement of the same formulas with Python floats and
    math.*, used to check the vectorised counter above; it shares no code
    with it.  Returns {(ir, i2): [DD, e+ sum, ex sum]} over in-domain pairs.
    """
    import math
    pos, shp, axis, e = prep["positions"], prep["shapes"], prep["axis"], prep["e"]
    w, ws = prep["weight"], prep["weight_shape"]
    los, nl = int(prep["los"]), [int(v) for v in prep["not_los"]]
    nr = len(r_bins) - 1
    n2 = len(second_bins) - 1
    half = boxsize / 2.0
    log_step = (np.log10(r_max) - np.log10(r_min)) / nr
    out = {}
    for p in range(len(pos)):
        for s in range(len(shp)):
            d = []
            for c in range(3):
                v = float(shp[s, c]) - float(pos[p, c])
                if periodic:
                    if v > half:
                        v -= boxsize
                    if v < -half:
                        v += boxsize
                d.append(v)
            a, b, l = d[nl[0]], d[nl[1]], d[los]
            rp = math.sqrt(a * a + b * b)
            if mode == "rppi":
                r = rp
                if not (r_bins[0] <= r < r_bins[-1] and second_bins[0] <= l < second_bins[-1]):
                    continue
                x2 = l
                step2 = (second_bins[-1] - second_bins[0]) / n2
            else:
                r = math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
                if not (rp > rp_cut and r_bins[0] <= r < r_bins[-1]):
                    continue
                x2 = l / r
                step2 = 2.0 / n2
            ir = int(np.floor(np.log10(r) / log_step - np.log10(r_bins[0]) / log_step))
            i2 = int(math.floor(x2 / step2 - second_bins[0] / step2))
            ir = nr - 1 if ir == nr else ir
            i2 = n2 - 1 if i2 == n2 else i2
            dot = 0.0 + (a / rp) * float(axis[s, 0])
            dot = dot + (b / rp) * float(axis[s, 1])
            if abs(dot) > 1.0:
                ep = ex = 0.0
            else:
                phi = math.acos(dot)
                ep = float(e[s]) * math.cos(2 * phi)
                ex = float(e[s]) * math.sin(2 * phi)
            acc = out.setdefault((ir, i2), [0.0, 0.0, 0.0])
            wp = float(w[p]) * float(ws[s])
            acc[0] += wp
            acc[1] += wp * ep
            acc[2] += wp * ex
    return out
