"""Expected lightcone outputs built from the oracle (test helper).

Restates, for one measure_xi_w / measure_xi_multipoles call of
MeasureIALightcone, the datasets the reference writes
(measure_IA.py:337-1060, measure_IA_base.py:670-743,
measure_jackknife.py:59-263), from oracle counts of every data/random
combination and every leave-one-patch-out realisation (masked recounts).
"""
import numpy as np

from oracle import lightcone as olc


def _fix(DD):
    DD = np.array(DD, dtype=np.float64)
    DD[DD == 0] = 1
    return DD


def _centres(edges):
    d = (edges[1:] - edges[:-1]) / 2.0
    return edges[:-1] + abs(d)


def _estimate(est, cnt, ns):
    """xi_g+ and xi_gg of one realisation (measure_IA_base.py:670-743, corr_type both)."""
    SpD, DD, SpR, SR, RD, RR = (cnt[k] for k in ("SpD", "DD", "SpR", "SR", "RD", "RR"))
    if est == "clusters":
        SRs = SR * (ns["D"] / ns["R_D"])
        xgp = SpD / DD - (SpR * (ns["D"] / ns["R_D"])) / SRs
        RDs = RD * (ns["S"] / ns["R_S"])
        RRs = RR * ((ns["S"] / ns["R_S"]) * (ns["D"] / ns["R_D"]))
        xgg = (DD - RDs - SRs) / RRs - 1
    else:
        RRs = RR * ((ns["S"] / ns["R_S"]) * (ns["D"] / ns["R_D"]))
        xgp = (SpD - SpR * (ns["D"] / ns["R_D"])) / RRs
        RDs = RD * (ns["S"] / ns["R_S"])
        SRs = SR * (ns["D"] / ns["R_D"])
        xgg = (DD - RDs - SRs) / RRs + 1
    return xgp, xgg


def _project(xi, mode, second):
    if mode == "rppi":
        d = second[1:] - second[:-1]
        return np.sum(xi * abs(d)[None, :], axis=1)
    mu = _centres(second)
    dmu = second[1:] - second[:-1]
    gp = np.sum((5 / 2.0) * (1 / 24.0) * 3.0 * (1 - mu * mu)[None, :] * xi["gp"] * dmu[None, :], axis=1)
    gg = np.sum(0.5 * xi["gg"] * dmu[None, :], axis=1)
    return gp, gg


def expected(mode, est, D, S, R, chi, r_bins, second, r_min, r_max, patches, rp_cut=None, h=1.0,
             over_h=False):
    """D, S, R: dicts of RA, DEC, w (+ e1, e2 for S); chi: dict of comoving
    distances per sample (already times h with over_h); patches: dict
    position/shape/randoms."""
    def cnt(pos, shp, pkey, skey, with_shapes, mp=None, ms=None):
        mp = slice(None) if mp is None else mp
        ms = slice(None) if ms is None else ms
        return olc.count(mode, pos["RA"][mp], pos["DEC"][mp], chi[pkey][mp], pos["w"][mp], shp["RA"][ms],
                         shp["DEC"][ms], chi[skey][ms], shp["w"][ms], r_bins, second, r_min, r_max,
                         e1=shp["e1"][ms] if with_shapes else None, e2=shp["e2"][ms] if with_shapes else None,
                         rp_cut=rp_cut, clamp=mode == "rppi", h=h, over_h=over_h)

    def all_counts(mD=None, mS=None, mR=None):
        a = cnt(D, S, "D", "S", True, mD, mS)
        b = cnt(R, S, "R", "S", True, mR, mS)
        c = cnt(D, R, "D", "R", False, mD, mR)
        d = cnt(R, R, "R", "R", False, mR, mR)
        return {"SpD": a["SplusD"][0], "SxD": a["ScrossD"][0], "DD": _fix(a["DD"][0]),
                "SpR": b["SplusD"][0], "SR": _fix(b["DD"][0]), "RD": _fix(c["DD"][0]), "RR": _fix(d["DD"][0])}

    ns = {"D": len(D["RA"]), "S": len(S["RA"]), "R_D": len(R["RA"]), "R_S": len(R["RA"])}
    main = all_counts()
    xgp, xgg = _estimate(est, main, ns)
    out = {"main": main, "xi_g_plus": xgp, "xi_gg": xgg}
    if mode == "rppi":
        out["w_g_plus"] = _project(xgp, mode, second)
        out["w_gg"] = _project(xgg, mode, second)
    else:
        out["multipoles_g_plus"], out["multipoles_gg"] = _project({"gp": xgp, "gg": xgg}, mode, second)
    pp, ps, pr = patches["position"], patches["shape"], patches["randoms"]
    lo, hi = int(min(ps)), int(max(ps))
    reals = {}
    for i in range(lo, hi + 1):
        c = all_counts(pp != i, ps != i, pr != i)
        nsi = {"D": int(np.sum(pp != i)), "S": int(np.sum(ps != i)), "R_D": int(np.sum(pr != i)),
               "R_S": int(np.sum(pr != i))}
        gp, gg = _estimate(est, c, nsi)
        if mode == "rppi":
            reals[i] = (c, _project(gp, mode, second), _project(gg, mode, second))
        else:
            mgp, mgg = _project({"gp": gp, "gg": gg}, mode, second)
            reals[i] = (c, mgp, mgg)
    out["reals"] = reals
    n = hi - lo + 1
    for k, name in ((1, "g_plus"), (2, "gg")):
        xs = [reals[i][k] for i in range(lo, hi + 1)]
        mean = np.mean(xs, axis=0)
        d = np.array(xs) - mean
        out[f"cov_{name}"] = (n - 1) / n * d.T @ d
        out[f"mean_{name}"] = mean
    return out
