"""Random configurations against the FP64 oracle (seeded, reproducible).

Hypothesis draws box sizes, sample sizes, binnings (mode, r range, bin
counts, pi_max or L/2), jackknife splits (none to 3^3), line of sight,
periodicity, weights and shapes-equal-positions samples; every draw is
counted through the C ABI and compared with oracle.brute: pair counts and
every tally equal, S+D, SxD and the S+D tallies within 1e-6 of
max(|ref|, sum|terms|), weighted DD within 1e-6.
"""
import numpy as np
import pytest
from hypothesis import HealthCheck, given, settings, strategies as st

from measure_ia_amd.synth import ellipticity_noise, shapes_from_components, wrap_box
from paircases import compare, engine_sample, oracle_counts, thresholds

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    from measure_ia_amd.engine import get_engine
    return get_engine(0)


@st.composite
def cases(draw):
    L = draw(st.floats(8.0, 150.0))
    r_min = draw(st.floats(0.05, 2.0))
    r_max = draw(st.floats(2.0 * r_min, max(2.01 * r_min, 0.9 * L)))
    cfg = dict(r_min=r_min, r_max=r_max, nr=draw(st.integers(1, 16)), n2=draw(st.integers(1, 20)),
               pi_max=draw(st.one_of(st.none(), st.floats(0.5, L / 2.0))),
               L_sub=draw(st.sampled_from([0, 1, 2, 3])), los=draw(st.sampled_from([0, 1, 2])),
               weights=draw(st.booleans()), periodic=draw(st.booleans()))
    n_pos = draw(st.integers(1, 700))
    auto = draw(st.booleans())
    n_shp = n_pos if auto else draw(st.integers(1, 300))
    mode = draw(st.sampled_from(["rppi", "rmu"]))
    seed = draw(st.integers(0, 2 ** 31 - 1))
    return L, cfg, n_pos, n_shp, auto, mode, seed


def _data(L, cfg, n_pos, n_shp, auto, seed):
    rng = np.random.default_rng(seed)
    pos = wrap_box(rng.uniform(0, L, (n_pos, 3)), L)
    shp = pos if auto else wrap_box(rng.uniform(0, L, (n_shp, 3)), L)
    axis, q = shapes_from_components(ellipticity_noise(rng, len(shp), 0.25))
    data = {"Position": pos, "Position_shape_sample": shp, "Axis_Direction": axis, "LOS": cfg["los"], "q": q}
    if cfg["weights"]:
        data["weight"] = rng.uniform(0.5, 1.5, len(pos))
        data["weight_shape_sample"] = rng.uniform(0.5, 1.5, len(shp))
    return data


@settings(max_examples=30, deadline=None, derandomize=True,
          suppress_health_check=[HealthCheck.function_scoped_fixture, HealthCheck.too_slow])
@given(case=cases())
def test_random_case(engine, case):
    L, cfg, n_pos, n_shp, auto, mode, seed = case
    print("case", L, cfg, n_pos, n_shp, auto, mode, seed, flush=True)
    data = _data(L, cfg, n_pos, n_shp, auto, seed)
    pr, ref, _ = oracle_counts(data, L, cfg, mode)
    p, sample = engine_sample(data, L, cfg)
    res = engine.count(sample, thresholds(cfg, L, mode), method="fast")
    K = cfg["L_sub"] ** 3 if cfg["L_sub"] else 0
    compare(res, ref, p.R, K, cfg["weights"])
    # one persistent wave (which also claims the tail chunks when its share
    # is large enough): the same counts
    one = engine.count(sample, thresholds(cfg, L, mode), method="fast", tuning={"waves": 1})
    np.testing.assert_array_equal(one.npairs, res.npairs)
    compare(one, ref, p.R, K, cfg["weights"])
