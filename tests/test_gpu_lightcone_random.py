"""Random lightcone configurations against the FP64 oracle (seeded, reproducible).

Hypothesis draws the sky window (RA width, DEC band up to |DEC| = 80 deg),
the comoving-distance range, sample sizes, mode, binning, pi_max, the
multipole r_p cut, over_h, patches, weights, counts only and shapes equal
to positions; every draw is counted through mia_count_lightcone and
compared with oracle/lightcone.py: pair counts and every patch tally equal,
DD within 1e-12 relative, S+D and SxD within 1e-6 of max(|ref|, sum|terms|).
"""
import numpy as np
import pytest
from hypothesis import HealthCheck, given, settings, strategies as st

from lccases import bins, device_arrays, oracle_counts
from measure_ia_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    from measure_ia_amd.engine import get_engine
    return get_engine(0)


@st.composite
def cases(draw):
    mode = draw(st.sampled_from(["rppi", "rmu"]))
    r_min = draw(st.floats(0.1, 2.0))
    cfg = dict(mode=mode, r_min=r_min, r_max=draw(st.floats(2.0 * r_min, 40.0)), nr=draw(st.integers(1, 12)),
               n2=draw(st.integers(1, 16)), pi_max=draw(st.floats(1.0, 60.0)), shapes=draw(st.booleans()),
               h=0.7, over_h=draw(st.booleans()),
               rp_cut=draw(st.one_of(st.none(), st.floats(0.1, 3.0))) if mode == "rmu" else None,
               P=draw(st.sampled_from([0, 2, 5])), weights=draw(st.booleans()), clamp=mode == "rppi")
    ra0 = draw(st.floats(0.0, 300.0))
    dec0 = draw(st.floats(-80.0, 60.0))
    chi0 = draw(st.floats(30.0, 1000.0))
    kw = dict(ra=(ra0, ra0 + draw(st.floats(1.0, 60.0))), dec=(dec0, min(dec0 + draw(st.floats(1.0, 30.0)), 80.0)),
              chi=(chi0, chi0 + draw(st.floats(20.0, 1500.0))))
    n_pos = draw(st.integers(1, 600))
    auto = draw(st.booleans())
    n_shape = n_pos if auto else draw(st.integers(1, 300))
    seed = draw(st.integers(0, 2 ** 31 - 1))
    return cfg, kw, n_pos, n_shape, auto, seed


def _arrays(cfg, kw, n_pos, n_shape, auto, seed):
    a = synth.lightcone(n_pos, n_shape, seed=seed, weights=cfg["weights"], **kw)
    if auto:
        for k in ("RA", "DEC", "chi"):
            a[f"{k}_shape_sample"] = a[k].copy()
        e12 = synth.ellipticity_noise(np.random.default_rng(seed), n_pos, 0.25)
        a["e1"], a["e2"] = e12[:, 0], e12[:, 1]
    if not cfg["weights"]:
        a["weight"] = np.ones(len(a["RA"]))
        a["weight_shape_sample"] = np.ones(len(a["RA_shape_sample"]))
    if cfg["over_h"]:
        a["chi"] = a["chi"] * cfg["h"]
        a["chi_shape_sample"] = a["chi_shape_sample"] * cfg["h"]
    if cfg["P"]:
        lo, hi = kw["ra"]
        a["patch"] = np.clip(((a["RA"] - lo) / (hi - lo) * cfg["P"]).astype(int), 0, cfg["P"] - 1)
        a["patch_s"] = np.clip(((a["RA_shape_sample"] - lo) / (hi - lo) * cfg["P"]).astype(int), 0,
                               cfg["P"] - 1)
    return a


def _rel(got, ref, scale):
    den = np.maximum(np.abs(ref), scale)
    d = np.abs(got - ref)
    return np.where(den > 0, d / np.where(den > 0, den, 1), d)


@settings(max_examples=30, deadline=None, derandomize=True,
          suppress_health_check=[HealthCheck.function_scoped_fixture, HealthCheck.too_slow])
@given(case=cases())
def test_random_lightcone_case(engine, case):
    cfg, kw, n_pos, n_shape, auto, seed = case
    print("case", cfg, kw, n_pos, n_shape, auto, seed, flush=True)
    a = _arrays(cfg, kw, n_pos, n_shape, auto, seed)
    _, _, th = bins(cfg)
    res = engine.count_lightcone(device_arrays(a, cfg), th)
    ref = oracle_counts(a, cfg)
    np.testing.assert_array_equal(res.npairs, ref["NP"])
    scale = max(np.abs(ref["DD"]).max(), 1e-300)
    assert np.abs(res.dd - ref["DD"]).max() <= 1e-12 * scale
    if cfg["shapes"]:
        assert _rel(res.splus, ref["SplusD"], ref["SplusD_abs"]).max() <= 1e-6
        assert _rel(res.scross, ref["ScrossD"], ref["ScrossD_abs"]).max() <= 1e-6
    # one persistent wave (which also claims the tail chunks when its share
    # is large enough): the same counts
    one = engine.count_lightcone(device_arrays(a, cfg), th, tuning={"waves": 1})
    np.testing.assert_array_equal(one.npairs, res.npairs)
    assert np.abs(one.dd - res.dd).max() <= 1e-12 * scale
