"""Host-side input preparation of a box measurement.

Every box counter of the reference starts with the same block
(src/measureia/measure_w_box_jk.py:97-155, measure_m_box_jk.py:102-163):
apply the masks, normalise Axis_Direction, turn q into |e|, compute the
responsivity R, pick the line-of-sight column.  That block is O(N) numpy
work and stays on the host; this module restates it and hands the arrays to
the device engine.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Prepared:
    positions: np.ndarray
    shapes: np.ndarray
    axis: np.ndarray
    e: np.ndarray
    weight: np.ndarray
    weight_shape: np.ndarray
    R: float
    los: int
    not_los: np.ndarray
    # the two sums of R, sum(w_s (1 - e^2/2)) and sum(w_s), kept for R_jk
    sum_wa: float = 0.0
    sum_w: float = 0.0

    @property
    def n_pos(self):
        return len(self.positions)

    @property
    def n_shape(self):
        return len(self.shapes)


def unit_weights(n: int) -> np.ndarray:
    """n weights of 1.0 as a read-only view of a single value."""
    return np.broadcast_to(np.float64(1.0), (n,))


def is_unit(w) -> bool:
    """Whether w is a view made by unit_weights (no scan needed)."""
    return isinstance(w, np.ndarray) and w.ndim == 1 and w.strides == (0,) and (w.size == 0 or w[0] == 1.0)


def sequential_sum(a: np.ndarray) -> float:
    """Python's builtin sum over an array (left to right, starting at 0)."""
    a = np.asarray(a, dtype=np.float64)
    if a.size == 0:
        return 0
    return float(np.cumsum(a)[-1])


def prepare(data: dict, masks=None, ellipticity: str = "distortion") -> Prepared:
    """Selection, |e|, R and the LOS split of one measurement.

    measure_w_box_jk.py:97-139: a mask dict selects rows of every array; when
    it lacks 'weight' / 'weight_shape_sample', the first k weights are taken
    (k = number of selected rows), and the default mask is stored back into
    ``masks`` as the reference does.  R = sum(w_s (1 - e^2/2)) / sum(w_s)
    with Python's sequential sum.
    """
    n_pos = len(data["Position"])
    n_shape = len(data["Position_shape_sample"])
    # absent weights are ones: read-only views of one value, which the engine
    # recognises without scanning them
    w_all = data["weight"] if "weight" in data else unit_weights(n_pos)
    ws_all = data["weight_shape_sample"] if "weight_shape_sample" in data else unit_weights(n_shape)
    if masks is None:
        pos = data["Position"]
        shp = data["Position_shape_sample"]
        axis_v = data["Axis_Direction"]
        q = data["q"]
        w, ws = w_all, ws_all
    else:
        pos = data["Position"][masks["Position"]]
        shp = data["Position_shape_sample"][masks["Position_shape_sample"]]
        axis_v = data["Axis_Direction"][masks["Axis_Direction"]]
        q = data["q"][masks["q"]]
        if "weight" not in masks:
            m = np.ones(n_pos, dtype=bool)
            m[sum(masks["Position"]):n_pos] = 0
            masks["weight"] = m
        if "weight_shape_sample" not in masks:
            m = np.ones(n_shape, dtype=bool)
            m[sum(masks["Position_shape_sample"]):n_shape] = 0
            masks["weight_shape_sample"] = m
        w = w_all[masks["weight"]]
        ws = ws_all[masks["weight_shape_sample"]]
    axis_v = np.asarray(axis_v, dtype=np.float64)
    axis = axis_v / np.sqrt(axis_v[:, 0] ** 2 + axis_v[:, 1] ** 2)[:, None] if len(axis_v) else axis_v.reshape(0, 2)
    q = np.asarray(q, dtype=np.float64)
    if ellipticity == "distortion":
        e = (1 - q ** 2) / (1 + q ** 2)
    elif ellipticity == "ellipticity":
        e = (1 - q) / (1 + q)
    else:
        raise ValueError("Invalid value for ellipticity. Choose 'distortion' or 'ellipticity'.")
    ws = np.asarray(ws, dtype=np.float64)
    sum_wa = sequential_sum(ws * (1 - e ** 2 / 2.0))
    sum_w = float(len(ws)) if is_unit(ws) else sequential_sum(ws)
    R = sum_wa / sum_w
    los = int(data["LOS"])
    not_los = np.array([c for c in range(3) if c != los])
    w = w if is_unit(w) else np.ascontiguousarray(w, dtype=np.float64)
    return Prepared(np.ascontiguousarray(pos, dtype=np.float64),
                    np.ascontiguousarray(shp, dtype=np.float64),
                    np.ascontiguousarray(axis, dtype=np.float64), np.ascontiguousarray(e),
                    w, ws, R, los, not_los, sum_wa, sum_w)


def jackknife_regions(xyz: np.ndarray, boxsize: float, L_subboxes: int) -> np.ndarray:
    """Region id per object, vectorised (measure_IA_base.py:404-452).

    Per axis the candidate index floor(x / L_sub) and its neighbours are
    tested with the reference's strict bounds i*L_sub < x < (i+1)*L_sub,
    the bounds computed as float(i) * L_sub like the reference's loop; an
    object that fails on any axis keeps region 0.
    """
    xyz = np.asarray(xyz, dtype=np.float64)
    L_sub = (boxsize / 2.0) * 2.0 / L_subboxes
    idx = np.full(xyz.shape, -1, dtype=np.int64)
    with np.errstate(invalid="ignore"):
        base = np.floor(xyz / L_sub)
    base = np.nan_to_num(base, nan=-10, posinf=-10, neginf=-10).astype(np.int64)
    for d in (-1, 0, 1):
        i = base + d
        ok = (i >= 0) & (i < L_subboxes)
        lo = i.astype(np.float64) * L_sub
        hi = (i + 1).astype(np.float64) * L_sub
        hit = ok & (xyz > lo) & (xyz < hi) & (idx < 0)
        idx = np.where(hit, i, idx)
    good = np.all(idx >= 0, axis=1)
    reg = idx[:, 0] * L_subboxes * L_subboxes + idx[:, 1] * L_subboxes + idx[:, 2]
    return np.where(good, reg, 0).astype(np.int64)
