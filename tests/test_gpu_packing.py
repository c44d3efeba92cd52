"""Work-item packing and planner edge cases against the FP64 oracle.

The planner packs the position ranges of up to 8 neighbouring stencil columns
of one (region, shift class) into one item (at most kItemRanges ranges and the
item's position limit, four open items per plan thread, a fifth class
finishing the fullest), and takes each block's slab runs from one table
(walked per column when a block has more than 15 runs).  These cases drive
those paths with tunings that the autotuner would not pick:

- few positions per item (ranges split across items, items of many ranges);
- fine grids and single-slab blocks (many ranges per column group, many keys);
- small shape blocks;
- 15 jackknife slabs along the line of sight inside the reach (more slab runs
  than the table holds, and more keys than the planner's LDS tables hold);
- periodic wraps, open boxes and objects on jackknife faces.

Every count is compared with oracle.brute as in the parity suite: pair
counts and tallies equal, sums within 1e-6.
"""
import numpy as np
import pytest

from measure_ia_amd._capi import MIA_FAST
from measure_ia_amd.synth import ellipticity_noise, shapes_from_components, wrap_box
from paircases import compare, engine_sample, make_case, oracle_counts, thresholds

pytestmark = pytest.mark.gpu

TUNINGS = [
    {"item_positions": 33},
    {"item_positions": 64, "block_shapes": 7},
    {"columns_per_side": 24, "slabs_per_side": 24, "slabs_per_block": 1},
    {"columns_per_side": 12, "slabs_per_side": 36, "slabs_per_block": 3, "item_positions": 100},
]


@pytest.fixture(scope="module")
def engine():
    from measure_ia_amd.engine import get_engine
    return get_engine(0)


def _fit(tu, cfg):
    """Round the grid overrides to multiples of the jackknife split."""
    n = cfg["L_sub"] or 1
    out = dict(tu)
    for k in ("columns_per_side", "slabs_per_side"):
        if k in out:
            out[k] = max(n, out[k] // n * n)
    return out


@pytest.mark.parametrize("mode", ["rppi", "rmu"])
@pytest.mark.parametrize("name", ["uniform_jk27", "clustered_jk8", "tiny_box_wrap", "open_box_jk27",
                                  "reach_beyond_half_box", "lattice_boundaries"])
def test_packing_tunings_match_oracle(engine, name, mode):
    data, L, cfg = make_case(name)
    pr, ref, _ = oracle_counts(data, L, cfg, mode)
    p, sample = engine_sample(data, L, cfg)
    th = thresholds(cfg, L, mode)
    K = cfg["L_sub"] ** 3 if cfg["L_sub"] else 0
    base = engine.count(sample, th, method="fast")
    compare(base, ref, p.R, K, cfg["weights"])
    for tu in TUNINGS:
        tu = _fit(tu, cfg)
        print(name, mode, tu, flush=True)
        res = engine.count(sample, th, method="fast", tuning=tu)
        np.testing.assert_array_equal(res.npairs, base.npairs)
        compare(res, ref, p.R, K, cfg["weights"])


@pytest.mark.parametrize("mode", ["rppi", "rmu"])
def test_many_line_of_sight_regions(engine, mode):
    # 15 regions per side: a block's reach along the line of sight (the whole
    # box) crosses 15 jackknife slabs plus the periodic half-box split, more
    # slab runs than a block's table holds; 3376^2 planner keys do not fit the
    # planner's LDS tables either
    rng = np.random.default_rng(1515)
    L = 30.0
    pos = wrap_box(rng.uniform(0, L, (500, 3)), L)
    shp = wrap_box(rng.uniform(0, L, (250, 3)), L)
    axis, q = shapes_from_components(ellipticity_noise(rng, len(shp), 0.25))
    cfg = dict(r_min=0.2, r_max=3.0, nr=5, n2=6, pi_max=None, L_sub=15, los=2, weights=False, periodic=True)
    data = {"Position": pos, "Position_shape_sample": shp, "Axis_Direction": axis, "LOS": 2, "q": q}
    pr, ref, _ = oracle_counts(data, L, cfg, mode)
    p, sample = engine_sample(data, L, cfg)
    th = thresholds(cfg, L, mode)
    K = 15 ** 3
    for tu in ({}, {"slabs_per_side": 30, "slabs_per_block": 1}, {"item_positions": 17}):
        print(mode, tu, flush=True)
        res = engine.count(sample, th, method="fast", tuning=tu)
        assert res.stats["method"] == MIA_FAST, res.stats
        compare(res, ref, p.R, K, cfg["weights"])
