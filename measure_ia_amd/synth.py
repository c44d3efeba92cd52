"""Synthetic galaxy catalogues for benchmarks and tests.

Produces the data dict MeasureIABox takes (Position, Position_shape_sample,
Axis_Direction, LOS, q, optional weights), following the input construction
of the benchmark configurations:

  * positions float64 in [0, L): x - L*floor(x/L), and x == L mapped to 0;
  * shapes from ellipticity components (e1, e2): theta = atan2(e2, e1)/2,
    Axis_Direction = (cos theta, sin theta) in the projected plane,
    |e| = sqrt(e1^2 + e2^2), q = sqrt((1-|e|)/(1+|e|)) -- the inverse of the
    default 'distortion' definition e = (1-q^2)/(1+q^2); noise draws are
    redrawn until |e| < 1;
  * clustered boxes: Gaussian clusters (centres uniform, members N(0, sigma^2)
    per axis, Poisson multiplicity) plus a uniform background;
  * aligned shapes: e = A (cos 2phi_c, sin 2phi_c) + noise, phi_c the
    minimum-image projected angle from the galaxy to its cluster centre.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np


def wrap_box(x, L):
    x = x - L * np.floor(x / L)
    x[x >= L] = 0.0
    return x


def ellipticity_noise(rng, n, sigma):
    e = rng.normal(0.0, sigma, (n, 2))
    bad = np.hypot(e[:, 0], e[:, 1]) >= 1.0
    while bad.any():
        e[bad] = rng.normal(0.0, sigma, (int(bad.sum()), 2))
        bad = np.hypot(e[:, 0], e[:, 1]) >= 1.0
    return e


def shapes_from_components(e12):
    theta = 0.5 * np.arctan2(e12[:, 1], e12[:, 0])
    axis = np.stack([np.cos(theta), np.sin(theta)], axis=1)
    emag = np.hypot(e12[:, 0], e12[:, 1])
    q = np.sqrt((1.0 - emag) / (1.0 + emag))
    return axis, q


@dataclass
class Catalogue:
    data: dict
    boxsize: float
    cluster_id: Optional[np.ndarray] = None   # per position, -1 for background


def uniform(n, L, seed, sigma_e=0.2, los=2):
    """n uniform galaxies, shape sample = position sample (configuration 0)."""
    rng = np.random.default_rng(seed)
    pos = wrap_box(rng.uniform(0.0, L, (n, 3)), L)
    axis, q = shapes_from_components(ellipticity_noise(rng, n, sigma_e))
    data = {"Position": pos, "Position_shape_sample": pos, "Axis_Direction": axis, "LOS": los,
            "q": q}
    return Catalogue(data, L)


def clustered_positions(rng, n_total, L, n_clusters, sigma, mean_members, n_background=None):
    centres = rng.uniform(0.0, L, (n_clusters, 3))
    counts = rng.poisson(mean_members, n_clusters)
    n_members = int(counts.sum())
    if n_background is None:
        n_background = max(0, n_total - n_members)
    cid = np.repeat(np.arange(n_clusters), counts)
    members = centres[cid] + rng.normal(0.0, sigma, (n_members, 3))
    back = rng.uniform(0.0, L, (n_background, 3))
    pos = wrap_box(np.concatenate([members, back]), L)
    cluster = np.concatenate([cid, np.full(n_background, -1)])
    return pos, cluster, centres


def aligned_shapes(rng, pos, cluster, centres, L, amplitude, sigma_e, los=2):
    n = len(pos)
    e = ellipticity_noise(rng, n, sigma_e)
    nl = [c for c in range(3) if c != los]
    m = cluster >= 0
    d = centres[cluster[m]][:, nl] - pos[m][:, nl]
    d = d - L * np.round(d / L)
    phi = np.arctan2(d[:, 1], d[:, 0])
    sig = np.stack([amplitude * np.cos(2 * phi), amplitude * np.sin(2 * phi)], axis=1)
    e[m] = e[m] + sig
    bad = np.hypot(e[:, 0], e[:, 1]) >= 1.0
    while bad.any():
        extra = ellipticity_noise(rng, int(bad.sum()), sigma_e)
        base = np.zeros_like(extra)
        idx = np.nonzero(bad)[0]
        mm = cluster[idx] >= 0
        if mm.any():
            dd = centres[cluster[idx[mm]]][:, nl] - pos[idx[mm]][:, nl]
            dd = dd - L * np.round(dd / L)
            ph = np.arctan2(dd[:, 1], dd[:, 0])
            base[mm] = np.stack([amplitude * np.cos(2 * ph), amplitude * np.sin(2 * ph)], axis=1)
        e[idx] = base + extra
        bad = np.hypot(e[:, 0], e[:, 1]) >= 1.0
    return shapes_from_components(e)


def clustered(n_total, n_shapes, L, n_clusters, sigma, mean_members, amplitude, sigma_e, seed,
              background_fraction=None, weights=False, los=2):
    """Clustered box with an aligned shape subset (configurations 1, 2, 4)."""
    rng = np.random.default_rng(seed)
    n_bg = None
    if background_fraction is not None:
        n_bg = int(round(background_fraction * n_total))
    pos, cluster, centres = clustered_positions(rng, n_total, L, n_clusters, sigma, mean_members, n_bg)
    idx = np.sort(rng.choice(len(pos), size=min(n_shapes, len(pos)), replace=False))
    axis, q = aligned_shapes(rng, pos[idx], cluster[idx], centres, L, amplitude, sigma_e, los)
    data = {"Position": pos, "Position_shape_sample": pos[idx].copy(), "Axis_Direction": axis,
            "LOS": los, "q": q}
    if weights:
        data["weight"] = rng.uniform(0.5, 1.5, len(pos))
        data["weight_shape_sample"] = rng.uniform(0.5, 1.5, len(idx))
    return Catalogue(data, L, cluster)


# The benchmark workloads (sizes as stated; seeds fixed per configuration).
def config(name: str, seed: Optional[int] = None, scale: float = 1.0) -> Catalogue:
    """Catalogue of a named workload; ``scale`` < 1 shrinks counts for tests."""
    if name == "uniform20k":
        return uniform(int(20000 * scale), 205.0, 0 if seed is None else seed)
    if name == "clustered1M":
        n = int(1_000_000 * scale)
        return clustered(n, int(200_000 * scale), 205.0, max(1, int(5000 * scale)), 1.0, 150,
                         0.1, 0.2, 1 if seed is None else seed)
    if name == "multipoles2M":
        n = int(2_000_000 * scale)
        return clustered(n, int(500_000 * scale), 205.0, max(1, int(10000 * scale)), 1.0, 150,
                         0.1, 0.2, 2 if seed is None else seed)
    if name == "large20M":
        n = int(20_000_000 * scale)
        return clustered(n, int(5_000_000 * scale), 500.0, max(1, int(20000 * scale)), 1.5, 700,
                         0.05, 0.25, 4 if seed is None else seed, background_fraction=0.3)
    raise KeyError(name)


def uneven_samples(n_samples=12, parent=3_000_000, L=205.0, lo=10_000, hi=800_000, shape_frac=0.3,
                   seed=3, scale=1.0):
    """Samples of log-uniform sizes cut from one clustered box, weights U(0.5, 1.5)."""
    rng = np.random.default_rng(seed)
    n_parent = int(parent * scale)
    pos, cluster, centres = clustered_positions(rng, n_parent, L, max(1, int(15000 * scale)), 1.0, 150)
    out = []
    for _ in range(n_samples):
        n = int(round(np.exp(rng.uniform(np.log(lo * scale), np.log(hi * scale)))))
        n = max(2, min(n, len(pos)))
        sel = np.sort(rng.choice(len(pos), size=n, replace=False))
        ns = max(1, int(round(shape_frac * n)))
        sidx = np.sort(rng.choice(sel, size=ns, replace=False))
        axis, q = aligned_shapes(rng, pos[sidx], cluster[sidx], centres, L, 0.1, 0.2)
        data = {"Position": pos[sel].copy(), "Position_shape_sample": pos[sidx].copy(),
                "Axis_Direction": axis, "LOS": 2, "q": q,
                "weight": rng.uniform(0.5, 1.5, n), "weight_shape_sample": rng.uniform(0.5, 1.5, ns)}
        out.append(Catalogue(data, L))
    return out


# ---------------------------------------------------------------------------
# lightcones

def _sky_to_cart(ra, dec, chi):
    r, d = np.radians(ra), np.radians(dec)
    return np.stack([chi * np.cos(d) * np.cos(r), chi * np.cos(d) * np.sin(r), chi * np.sin(d)], 1)


def _cart_to_sky(x):
    chi = np.sqrt(np.sum(x * x, axis=1))
    dec = np.degrees(np.arcsin(np.clip(x[:, 2] / chi, -1.0, 1.0)))
    ra = np.degrees(np.arctan2(x[:, 1], x[:, 0])) % 360.0
    return ra, dec, chi


def lightcone_points(rng, n, ra=(10.0, 40.0), dec=(-10.0, 10.0), chi=(200.0, 1400.0),
                     clustered_frac=0.5, n_clusters=None, sigma=1.5):
    """n objects in a lightcone wedge, uniform in comoving volume, a fraction
    of them in Gaussian clusters (comoving sigma) around uniform centres.
    Returns ra, dec (degrees) and comoving distance chi."""
    def uniform_pts(m):
        r = rng.uniform(ra[0], ra[1], m)
        s0, s1 = np.sin(np.radians(dec[0])), np.sin(np.radians(dec[1]))
        d = np.degrees(np.arcsin(rng.uniform(s0, s1, m)))
        c = np.cbrt(rng.uniform(chi[0] ** 3, chi[1] ** 3, m))
        return r, d, c

    n_cl = int(round(n * clustered_frac))
    r, d, c = uniform_pts(n - n_cl)
    if n_cl > 0:
        k = n_clusters or max(1, n_cl // 50)
        cr, cd, cc = uniform_pts(k)
        centres = _sky_to_cart(cr, cd, cc)
        members = rng.integers(0, k, n_cl)
        x = centres[members] + rng.normal(0.0, sigma, (n_cl, 3))
        mr, md, mc = _cart_to_sky(x)
        # keep RA continuous across the wedge (no jump from 0 to 360)
        mid = 0.5 * (ra[0] + ra[1])
        mr = (mr - mid + 180.0) % 360.0 - 180.0 + mid
        # keep members inside the wedge's chi range
        mc = np.clip(mc, chi[0], chi[1])
        r, d, c = np.concatenate([r, mr]), np.concatenate([d, md]), np.concatenate([c, mc])
    p = rng.permutation(n)
    return r[p], d[p], c[p]


def lightcone(n_pos, n_shape, seed=0, sigma_e=0.25, weights=False, **kw):
    """Position and shape samples of one lightcone wedge (shapes a random
    subset of the positions) with ellipticity components e1, e2.  Returns a
    dict of arrays keyed like the reference's lightcone data dict, with the
    comoving distances under "chi" / "chi_shape_sample" (Redshift is filled
    by the caller from a cosmology when needed)."""
    rng = np.random.default_rng(seed)
    ra, dec, chi = lightcone_points(rng, n_pos, **kw)
    sel = rng.choice(n_pos, n_shape, replace=n_shape > n_pos)
    e12 = ellipticity_noise(rng, n_shape, sigma_e)
    out = {"RA": ra, "DEC": dec, "chi": chi, "RA_shape_sample": ra[sel], "DEC_shape_sample": dec[sel],
           "chi_shape_sample": chi[sel], "e1": e12[:, 0], "e2": e12[:, 1]}
    if weights:
        out["weight"] = rng.uniform(0.5, 1.5, n_pos)
        out["weight_shape_sample"] = rng.uniform(0.5, 1.5, n_shape)
    return out


def redshift_of_chi(cosmology, chi):
    """Inverse of the cosmology's chi(a): z with chi(1/(1+z)) = chi (table inversion)."""
    z = np.linspace(0.0, 10.0, 200001)
    c = np.asarray(cosmology.comoving_radial_distance(1.0 / (1.0 + z)))
    return np.interp(np.asarray(chi), c, z)
