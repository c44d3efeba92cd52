"""Brute-force FP64 lightcone pair counters (oracle) -- TEST INFRASTRUCTURE ONLY.

Restates the per-position loops of the reference's lightcone counters, with
the reference's own arithmetic in its own operation order:

  count_rppi  _measure_xi_rp_pi_lightcone_brute     (src/measureia/measure_w_lightcone.py:45-214)
              _count_pairs_xi_rp_pi_lightcone_brute (measure_w_lightcone.py:216-343)
  count_rmu   _measure_xi_r_mur_lightcone_brute     (src/measureia/measure_m_lightcone.py:45-219)
              _count_pairs_xi_r_mur_lightcone_brute (measure_m_lightcone.py:221-360)
  realisations  the leave-one-patch-out reruns of
              _measure_jackknife_realisations_lightcone (src/measureia/measure_jackknife.py:59-171)

The inputs are the arrays the reference selects and derives before its loop:
RA, DEC (degrees), the comoving distances chi (times h when over_h, as the
reference scales LOS_all), weights, and e1, e2 for the shape sample.  The
distance computation itself (pyccl) is outside this path: callers pass chi.

The loop runs over blocks of positions and is vectorised over all shapes,
so every per-pair operation is the reference's; histograms are accumulated
with np.bincount (pair counts exact, floating sums equal to np.add.at's up to
the order of FP64 additions).  Returned sums are raw: the reference's
DD[DD == 0] = 1 and S+D/DD are host post-processing (post_process()).
"""
import numpy as np

_BLOCK_ELEMS = 1 << 21


def shape_angles(e1, e2):
    """e and phi_axis as the reference derives them (measure_w_lightcone.py:134-139)."""
    theta = 1.0 / 2 * np.arctan2(e2, e1)
    axis = np.array([np.cos(theta), np.sin(theta)])
    axis_len = np.sqrt(np.sum(axis ** 2, axis=0))
    axis = axis / axis_len
    e = np.sqrt(e1 ** 2 + e2 ** 2)
    phi_axis = np.arctan2(axis[1], axis[0])
    return e, phi_axis


def _index(ind, n, clamp):
    """int cast, the clamp index == n -> n-1 where the reference has it, numpy's negative wrap."""
    ind = np.array(ind, dtype=int)
    if clamp and np.any(ind == n):
        ind[ind == n] = n - 1
    ind = np.where(ind < 0, ind + n, ind)
    if ind.size and (ind.min() < 0 or ind.max() >= n):
        raise IndexError("bin index out of range (the reference raises here too)")
    return ind


def _separations(RA_n, DEC_n, chi_n, RA_s, DEC_s, chi_s):
    """pi, dx, dy of a block of positions (rows) against all shapes (columns).

    measure_w_lightcone.py:143-147: LOS = chi_s - chi_n,
    dx = ((RA_s - RA_n) / 180 * pi) * chi_n * cos(DEC_n / 180 * pi),
    dy = ((DEC_s - DEC_n) / 180 * pi) * chi_n.
    """
    LOS = chi_s[None, :] - chi_n[:, None]
    dra = (RA_s[None, :] - RA_n[:, None]) / 180 * np.pi
    ddec = (DEC_s[None, :] - DEC_n[:, None]) / 180 * np.pi
    dx = dra * chi_n[:, None] * np.cos(DEC_n / 180 * np.pi)[:, None]
    dy = ddec * chi_n[:, None]
    return LOS, dx, dy


def _tally(out, name, flat, weights, nb, P, pn, ps):
    """Add weights to the total and, with patches, to tally[pn] and tally[ps != pn]."""
    out[name][0] += np.bincount(flat, weights=weights, minlength=nb)
    if P:
        out[name][1:] += np.bincount(pn * nb + flat, weights=weights, minlength=P * nb).reshape(P, nb)
        o = ps != pn
        out[name][1:] += np.bincount((ps * nb + flat)[o], weights=None if weights is None else weights[o],
                                     minlength=P * nb).reshape(P, nb)


def count(mode, RA, DEC, chi, w, RA_s, DEC_s, chi_s, w_s, r_bins, second_bins, r_min, r_max, *,
          e1=None, e2=None, h=1.0, over_h=False, rp_cut=None, clamp=True, patch=None, patch_s=None,
          n_patches=0):
    """One lightcone count, 'rppi' or 'rmu'.

    Returns dict with (1+P, nr, n2) arrays NP (integer counts), DD, SplusD,
    ScrossD (raw sums, e+ and ex with the reference's sign), SplusD_abs /
    ScrossD_abs (sums of |terms|, the scale of the rounding tolerance) --
    [0] all pairs, [1+i] the tally of pairs touching patch i.  clamp: the
    index == n -> n-1 clamp of _measure_xi_rp_pi_lightcone_brute; the other
    three counters have none (they raise IndexError where it would apply).
    """
    RA, DEC, chi, w = (np.asarray(a, dtype=np.float64) for a in (RA, DEC, chi, w))
    RA_s, DEC_s, chi_s, w_s = (np.asarray(a, dtype=np.float64) for a in (RA_s, DEC_s, chi_s, w_s))
    nr = len(r_bins) - 1
    n2 = len(second_bins) - 1
    nb = nr * n2
    P = int(n_patches)
    shapes = e1 is not None
    log_step = (np.log10(r_max) - np.log10(r_min)) / nr
    if mode == "rppi":
        step2 = (second_bins[-1] - second_bins[0]) / n2
    else:
        step2 = 2.0 / n2
        if rp_cut is None:
            rp_cut = 0.0
    if shapes:
        e, phi_axis = shape_angles(np.asarray(e1, dtype=np.float64), np.asarray(e2, dtype=np.float64))
    out = {k: np.zeros((1 + P, nb)) for k in ("DD", "SplusD", "ScrossD", "SplusD_abs", "ScrossD_abs")}
    out["NP"] = np.zeros((1 + P, nb))
    pa = np.asarray(patch) if P else None
    ps_all = np.asarray(patch_s) if P else None
    Ns = len(RA_s)
    B = max(1, _BLOCK_ELEMS // max(Ns, 1))
    for b0 in range(0, len(RA), B):
        sl_ = slice(b0, b0 + B)
        LOS, dx, dy = _separations(RA[sl_], DEC[sl_], chi[sl_], RA_s, DEC_s, chi_s)
        pdx, pdy = dx, dy
        if over_h:
            pdx = dx * h
            pdy = dy * h
        rp = np.sqrt(pdx ** 2 + pdy ** 2)
        if mode == "rppi":
            sep_len = rp
            mask = (rp >= r_bins[0]) * (rp < r_bins[-1]) * (LOS >= second_bins[0]) * (LOS < second_bins[-1])
        else:
            sep_len = np.sqrt((dx ** 2 + dy ** 2) + LOS ** 2)
            with np.errstate(invalid="ignore", divide="ignore"):
                mu = LOS / sep_len
            mask = (rp > rp_cut) * (sep_len >= r_bins[0]) * (sep_len < r_bins[-1])
        if not mask.any():
            continue
        ip, iS = np.nonzero(mask)
        ind_r = _index(np.floor(np.log10(sep_len[mask]) / log_step - np.log10(r_bins[0]) / log_step),
                       nr, clamp)
        if mode == "rppi":
            ind_2 = np.floor(LOS[mask] / step2 - second_bins[0] / step2)
        else:
            ind_2 = np.floor(mu[mask] / step2 - second_bins[0] / step2)
        ind_2 = _index(ind_2, n2, clamp)
        flat = ind_r * n2 + ind_2
        wprod = w[b0 + ip] * w_s[iS]
        pn = pa[b0 + ip] if P else None
        ps = ps_all[iS] if P else None
        _tally(out, "NP", flat, None, nb, P, pn, ps)
        _tally(out, "DD", flat, wprod, nb, P, pn, ps)
        if shapes:
            with np.errstate(invalid="ignore", divide="ignore"):
                sdx = pdx[mask] / rp[mask]
                sdy = pdy[mask] / rp[mask]
                phi_sep = np.arctan2(sdy, sdx)
                phi = phi_axis[iS] - phi_sep
            e_plus = -e[iS] * np.cos(2 * phi)
            e_cross = -e[iS] * np.sin(2 * phi)
            e_plus[np.isnan(e_plus)] = 0.0
            e_cross[np.isnan(e_cross)] = 0.0
            sp = w[b0 + ip] * w_s[iS] * e_plus
            sx = w[b0 + ip] * w_s[iS] * e_cross
            _tally(out, "SplusD", flat, sp, nb, P, pn, ps)
            _tally(out, "ScrossD", flat, sx, nb, P, pn, ps)
            _tally(out, "SplusD_abs", flat, np.abs(sp), nb, P, pn, ps)
            _tally(out, "ScrossD_abs", flat, np.abs(sx), nb, P, pn, ps)
    res = {k: v.reshape(1 + P, nr, n2) for k, v in out.items()}
    res["NP"] = res["NP"].astype(np.int64)
    return res


def realisations(mode, arrays, patch, patch_s, patches, r_bins, second_bins, r_min, r_max, **kw):
    """Leave-one-patch-out counts by masking and recounting, as the reference
    does (measure_jackknife.py:84-130): realisation i drops every position and
    shape of patch i.  arrays = (RA, DEC, chi, w, RA_s, DEC_s, chi_s, w_s) and
    optional e1, e2 in kw.  Returns {i: count dict (no patches)}."""
    RA, DEC, chi, w, RA_s, DEC_s, chi_s, w_s = arrays
    e1 = kw.pop("e1", None)
    e2 = kw.pop("e2", None)
    out = {}
    for i in patches:
        mp = np.asarray(patch) != i
        ms = np.asarray(patch_s) != i
        out[i] = count(mode, RA[mp], DEC[mp], chi[mp], w[mp], RA_s[ms], DEC_s[ms], chi_s[ms], w_s[ms],
                       r_bins, second_bins, r_min, r_max,
                       e1=None if e1 is None else e1[ms], e2=None if e2 is None else e2[ms], **kw)
    return out


def post_process(SplusD, DD):
    """DD[DD == 0] = 1 and correlation = S+D/DD (measure_w_lightcone.py:197-199)."""
    DD = np.array(DD, dtype=np.float64)
    DD[np.where(DD == 0)] = 1
    return SplusD / DD, DD
