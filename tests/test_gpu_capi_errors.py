"""Error paths of the C ABI on a device: every invalid input is refused with
the documented MIA_E_* code and a message, and the context stays usable."""
import numpy as np
import pytest

from measure_ia_amd import binning
from measure_ia_amd._capi import MiaError
from measure_ia_amd.engine import SampleArrays

pytestmark = pytest.mark.gpu

L = 50.0


@pytest.fixture(scope="module")
def engine():
    from measure_ia_amd.engine import get_engine
    return get_engine(0)


def _th():
    r_bins = np.logspace(-1, 1, 6)
    return binning.rppi_thresholds(r_bins, np.linspace(-10, 10, 5), 0.1, 10.0)


def _sample(pos, shp, jk=2, **kw):
    rng = np.random.default_rng(0)
    ax = rng.normal(size=(len(shp), 2))
    ax /= np.linalg.norm(ax, axis=1)[:, None]
    return SampleArrays(pos, shp, ax, rng.uniform(0, 0.5, len(shp)), None, None, L, 2, jk,
                        (L / 2.0) * 2.0 / jk if jk else 0.0, **kw)


def _good():
    rng = np.random.default_rng(1)
    return rng.uniform(0, L, (500, 3)), rng.uniform(0, L, (200, 3))


@pytest.mark.parametrize("bad", [-1e-9, L + 1e-9, np.nan])
def test_out_of_box_coordinates_are_refused(engine, bad):
    pos, shp = _good()
    pos[17, 1] = bad
    with pytest.raises(MiaError) as e:
        engine.count(_sample(pos, shp), _th())
    assert e.value.code == -4 and "boxsize" in str(e.value)
    shp[3, 2] = bad
    with pytest.raises(MiaError) as e:
        engine.count(_sample(_good()[0], shp), _th())
    assert e.value.code == -4
    # the context still works after the error
    ok = engine.count(_sample(*_good()), _th())
    assert ok.npairs[0].sum() > 0


def test_bad_arguments_are_refused(engine):
    pos, shp = _good()
    with pytest.raises(MiaError) as e:
        engine.count(_sample(pos, shp, jk=65), _th())
    assert e.value.code == -1
    th = _th()
    th.r2_edges = th.r2_edges[::-1].copy()  # unsorted thresholds
    with pytest.raises(MiaError) as e:
        engine.count(_sample(pos, shp), th)
    assert e.value.code == -1 and "sorted" in str(e.value)
    big = binning.rppi_thresholds(np.logspace(-1, 1, 301), np.linspace(-10, 10, 5), 0.1, 10.0)
    with pytest.raises(MiaError) as e:
        engine.count(_sample(pos, shp), big)
    assert e.value.code == -1
    ok = engine.count(_sample(pos, shp), _th())
    assert ok.npairs[0].sum() > 0
    # region ids: a split whose ids would overflow int32, and no split at all
    for jk in (1025, 0):
        with pytest.raises(MiaError) as e:
            engine.region_ids(pos, jk, L / max(jk, 1))
        assert e.value.code == -1
    np.testing.assert_array_equal(engine.region_ids(pos[:3], 1, L), [0, 0, 0])


def test_empty_samples_give_zero_counts(engine):
    pos, shp = _good()
    for p_, s_ in ((pos, shp[:0]), (pos[:0], shp), (pos[:0], shp[:0])):
        r = engine.count(_sample(p_, s_), _th())
        assert r.npairs.sum() == 0 and np.all(r.splus == 0) and np.all(r.scross == 0)
    for method in ("plain", "fast"):
        r = engine.count(_sample(pos[:0], shp), _th(), method=method)
        assert r.npairs.sum() == 0
