"""Exact bin thresholds for the device counters.

The reference bins a pair with numpy FP64 arithmetic
(src/measureia/measure_w_box_jk.py:142-143, 178-193;
src/measureia/measure_m_box_jk.py:149-150, 187-204):

    in domain   r_bins[0] <= r < r_bins[-1]  (and pi_bins[0] <= pi < pi_bins[-1] for w)
    ind_r     = floor(np.log10(r)/D - np.log10(r_bins[0])/D),  D = (log10 rmax - log10 rmin)/nr
    ind_pi    = floor(pi/Dpi - pi_bins[0]/Dpi),                Dpi = (pi_bins[-1]-pi_bins[0])/npi
    ind_mu    = floor(mu/Dmu - mu_r_bins[0]/Dmu),              Dmu = 2.0/nmu
    index == n -> n-1; a negative index wraps (numpy indexing)

with r = sqrt(s), s the FP64 squared separation.  Every index is a
non-decreasing step function of one double (s, pi or mu), so it is fully
described by the smallest double at which it reaches each value.  Those
thresholds are bisected here over the ordered bit patterns of doubles, using
this host's own numpy (the same library the reference would use), and the
device then only compares exact FP64 values against them: no log10 on the
GPU, and the result matches numpy's formula bit-for-bit.

A self-check evaluates the formula on +-2048 ulps around every threshold and
raises if the step function is not reproduced (which would mean np.log10 is
not monotone here).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

_SIGN = np.int64(-0x8000000000000000)
_MAG = np.int64(0x7FFFFFFFFFFFFFFF)


def to_ordered(x) -> np.ndarray:
    """Map doubles to int64 keys with the same order (-0.0 and +0.0 -> 0)."""
    i = np.asarray(x, dtype=np.float64).view(np.int64)
    return np.where(i >= 0, i, -(i & _MAG))


def from_ordered(k) -> np.ndarray:
    k = np.asarray(k, dtype=np.int64)
    bits = np.where(k >= 0, k, (-k) | _SIGN)
    return bits.view(np.float64)


def _bisect_first_true(pred, lo: float, hi: float, targets: np.ndarray) -> np.ndarray:
    """Smallest double x in [lo, hi] with pred(x, t) for each target t.

    pred(xs, ts) -> bool array, monotone (False ... True) in x for every t;
    pred(hi, t) must hold.  Returns an array of doubles, one per target.
    """
    targets = np.asarray(targets)
    klo = np.full(targets.shape, to_ordered(lo), dtype=np.int64)
    khi = np.full(targets.shape, to_ordered(hi), dtype=np.int64)
    # invariant: pred(from(khi)) true; answer in (klo-1, khi]
    at_lo = pred(from_ordered(klo), targets)
    ans_lo = at_lo.copy()
    klo = np.where(at_lo, khi, klo)
    while True:
        # the key span can exceed int64 (negative to positive doubles)
        span = khi.view(np.uint64) - klo.view(np.uint64)
        active = span > 1
        if not active.any():
            break
        mid = klo + (span // np.uint64(2)).astype(np.int64)
        pm = pred(from_ordered(mid), targets)
        khi = np.where(active & pm, mid, khi)
        klo = np.where(active & ~pm, mid, klo)
    out = from_ordered(khi)
    return np.where(ans_lo, lo, out)


def _prev(x):
    return from_ordered(to_ordered(x) - 1)


def _next(x):
    return from_ordered(to_ordered(x) + 1)


def r_index(s, r_bins, r_min, r_max):
    """The reference's raw log bin index of squared separation(s) s (no clamp)."""
    nr = len(r_bins) - 1
    step = (np.log10(r_max) - np.log10(r_min)) / nr
    r = np.sqrt(np.asarray(s, dtype=np.float64))
    return np.floor(np.log10(r) / step - np.log10(r_bins[0]) / step)


def linear_index(x, lo_edge, step):
    """floor(x/step - lo_edge/step), the pi and mu index of the reference."""
    x = np.asarray(x, dtype=np.float64)
    return np.floor(x / step - lo_edge / step)


@dataclass
class Thresholds:
    mode: str                 # 'rppi' or 'rmu'
    r2_edges: np.ndarray      # [S_lo, T_1..T_{nr-1}, S_hi]
    x2_edges: np.ndarray      # rppi: [pi_lo, P_1..P_{n-1}, pi_hi]; rmu: [M_0..M_{n-1}]

    @property
    def n_r(self):
        return len(self.r2_edges) - 1

    @property
    def n_2(self):
        return len(self.x2_edges) - 1 if self.mode == "rppi" else len(self.x2_edges)


def _r2_thresholds(r_bins, r_min, r_max):
    nr = len(r_bins) - 1
    lo_r, hi_r = r_bins[0], r_bins[-1]
    big = float(hi_r) ** 2 * 4 + 1.0
    s_lo = _bisect_first_true(lambda s, t: np.sqrt(s) >= lo_r, 0.0, big, np.zeros(1))[0]
    s_hi = _bisect_first_true(lambda s, t: np.sqrt(s) >= hi_r, 0.0, big, np.zeros(1))[0]
    ks = np.arange(1, nr, dtype=np.float64)
    if nr > 1:
        T = _bisect_first_true(lambda s, t: r_index(s, r_bins, r_min, r_max) >= t, s_lo,
                               _prev(s_hi), ks)
    else:
        T = np.zeros(0)
    edges = np.concatenate([[s_lo], T, [s_hi]])
    ind0 = r_index(s_lo, r_bins, r_min, r_max)
    if ind0 != 0:
        raise ValueError(f"log bin index at the lower edge is {ind0}, expected 0")
    return edges


def _check_steps(edges_inner, index_fn, clamp_n, lo_dom, hi_dom):
    """Evaluate index_fn around every threshold and compare with the count rule."""
    pts = []
    for t in list(edges_inner) + [lo_dom, _prev(hi_dom)]:
        k = to_ordered(t) + np.arange(-2048, 2049)
        pts.append(from_ordered(k))
    if not pts:
        return
    x = np.concatenate(pts)
    x = x[(x >= lo_dom) & (x < hi_dom)]
    ind = np.array(index_fn(x), dtype=int)
    ind = np.where(ind >= clamp_n, clamp_n - 1, ind)
    cnt = np.searchsorted(np.asarray(edges_inner), x, side="right")
    if not np.array_equal(ind, np.minimum(cnt, clamp_n - 1)):
        raise ValueError("bin thresholds do not reproduce the reference index (non-monotone log10?)")


def rppi_thresholds(r_bins, pi_bins, r_min, r_max, check=True) -> Thresholds:
    """Thresholds of the (r_p, pi) grid of measure_xi_w."""
    r_bins = np.asarray(r_bins, dtype=np.float64)
    pi_bins = np.asarray(pi_bins, dtype=np.float64)
    r2 = _r2_thresholds(r_bins, r_min, r_max)
    npi = len(pi_bins) - 1
    step = (pi_bins[-1] - pi_bins[0]) / npi
    ks = np.arange(1, npi, dtype=np.float64)
    if npi > 1:
        P = _bisect_first_true(lambda x, t: linear_index(x, pi_bins[0], step) >= t, pi_bins[0],
                               _prev(pi_bins[-1]), ks)
    else:
        P = np.zeros(0)
    x2 = np.concatenate([[pi_bins[0]], P, [pi_bins[-1]]])
    if linear_index(pi_bins[0], pi_bins[0], step) != 0:
        raise ValueError("pi index at the lower edge is not 0")
    if check:
        nr = len(r_bins) - 1
        _check_steps(r2[1:-1], lambda s: r_index(s, r_bins, r_min, r_max), nr, r2[0], r2[-1])
        _check_steps(P, lambda x: linear_index(x, pi_bins[0], step), npi, pi_bins[0], pi_bins[-1])
    return Thresholds("rppi", r2, x2)


def rmu_thresholds(r_bins, mu_bins, r_min, r_max, check=True) -> Thresholds:
    """Thresholds of the (r, mu_r) grid of measure_xi_multipoles."""
    r_bins = np.asarray(r_bins, dtype=np.float64)
    mu_bins = np.asarray(mu_bins, dtype=np.float64)
    r2 = _r2_thresholds(r_bins, r_min, r_max)
    nmu = len(mu_bins) - 1
    step = 2.0 / nmu
    ks = np.arange(0, nmu, dtype=np.float64)
    M = _bisect_first_true(lambda x, t: linear_index(x, mu_bins[0], step) >= t, -4.0, 4.0, ks)
    if check:
        nr = len(r_bins) - 1
        _check_steps(r2[1:-1], lambda s: r_index(s, r_bins, r_min, r_max), nr, r2[0], r2[-1])
        # in [M_0, +inf): index = count of M_1.. below x, clamped
        _check_steps(M[1:], lambda x: linear_index(x, mu_bins[0], step), nmu, M[0], 4.0)
        below = _prev(M[0])
        if not linear_index(below, mu_bins[0], step) < 0:
            raise ValueError("mu index below M_0 is not negative")
    return Thresholds("rmu", r2, M)


def bin_rppi(s, pi, th: Thresholds):
    """Reference binning through the thresholds (host helper for tests)."""
    s = np.asarray(s)
    pi = np.asarray(pi)
    r2, x2 = th.r2_edges, th.x2_edges
    dom = (s >= r2[0]) & (s < r2[-1]) & (pi >= x2[0]) & (pi < x2[-1])
    ir = np.searchsorted(r2[1:-1], s, side="right")
    ip = np.searchsorted(x2[1:-1], pi, side="right")
    return dom, ir, ip


def bin_rmu(r2v, mu, th: Thresholds):
    r2, M = th.r2_edges, th.x2_edges
    n = len(M)
    dom = (r2v >= r2[0]) & (r2v < r2[-1])
    ir = np.searchsorted(r2[1:-1], r2v, side="right")
    im = np.where(mu < M[0], n - 1, np.searchsorted(M[1:], mu, side="right"))
    return dom, ir, im
