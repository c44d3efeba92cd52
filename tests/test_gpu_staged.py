"""The staged count paths through the C ABI: prepared samples
(mia_prepare_sample / mia_count_prepared) and the hipGraph replay of
mia_count_batch, each against the plain-launch count of the same inputs.

Integer pair counts and tallies must be equal bit for bit; the FP64 sums may
differ only by the rounding of their atomic order (1e-12 of the bin scale).
A graph is replayed several times with other eager counts and host syncs in
between, and with fresh input and output buffers, so a replay that reused a
stale pointer or a stale zero-fill would show.
"""
import numpy as np
import pytest

from measure_ia_amd import binning, hostprep
from measure_ia_amd.engine import SampleArrays
from measure_ia_amd.synth import uneven_samples
from oracle import brute, prep
from paircases import compare

pytestmark = pytest.mark.gpu

L = 60.0
R_MIN, R_MAX, NR, NPI, PI_MAX = 0.1, 10.0, 10, 8, 20.0


@pytest.fixture(scope="module")
def engine():
    from measure_ia_amd.engine import get_engine
    return get_engine(0)


def _th(mode):
    r_bins, pi_bins, mu_bins = prep.bin_edges(R_MIN, R_MAX, NR, NPI, PI_MAX, L)
    if mode == "rppi":
        return binning.rppi_thresholds(r_bins, pi_bins, R_MIN, R_MAX), (r_bins, pi_bins)
    return binning.rmu_thresholds(r_bins, mu_bins, R_MIN, R_MAX), (r_bins, mu_bins)


def _samples(seed=21):
    cats = uneven_samples(n_samples=5, parent=20000, L=L, lo=300, hi=6000, seed=seed)
    specs = [(3, 2), (3, 2), (2, 2), (0, 2), (3, 0)]
    out = []
    for cat, (jk, los) in zip(cats, specs):
        d = dict(cat.data)
        d["LOS"] = los
        p = hostprep.prepare(d)
        out.append((d, p, SampleArrays(positions=p.positions, shapes=p.shapes, axis=p.axis, e=p.e,
                                       weight=p.weight, weight_shape=p.weight_shape, boxsize=L,
                                       los=p.los, jk_per_side=jk,
                                       jk_sub=(L / 2.0) * 2.0 / jk if jk else 0.0)))
    return out


def _same(a, b, tag):
    np.testing.assert_array_equal(a.npairs, b.npairs, err_msg=tag)
    for name in ("dd", "splus", "scross"):
        x, y = getattr(a, name), getattr(b, name)
        assert np.abs(x - y).max() <= 1e-12 * max(np.abs(y).max(), 1e-300), (tag, name)
    if a.region_stats is not None:
        np.testing.assert_array_equal(a.region_stats[:, :2], b.region_stats[:, :2], err_msg=tag)
        np.testing.assert_allclose(a.region_stats[:, 2:], b.region_stats[:, 2:], rtol=1e-13, err_msg=tag)


@pytest.mark.parametrize("mode", ["rppi", "rmu"])
def test_prepared_equals_unprepared(engine, mode):
    th, _ = _th(mode)
    sm = _samples()
    handles = [engine.prepare(s, th) for _, _, s in sm]
    # the prepared batch in one launch, and each handle alone
    many = engine.count_prepared(handles, th)
    for (d, p, s), h, m in zip(sm, handles, many):
        ref = engine.count(s, th, method="fast")
        _same(m, ref, "batch of handles")
        one = engine.count_prepared([h], th)[0]
        _same(one, ref, "single handle")
        assert one.stats["ms_prepare"] < 0.05  # no sort on a prepared sample
    # a handle sorted for one binning serves another (the grid stays valid)
    other, _ = _th("rmu" if mode == "rppi" else "rppi")
    for (d, p, s), h in zip(sm[:2], handles[:2]):
        _same(engine.count_prepared([h], other)[0], engine.count(s, other, method="fast"), "other binning")
    for h in handles:
        h.release()


def test_prepared_against_oracle(engine):
    th, (r_bins, sb) = _th("rppi")
    d, p, s = _samples()[1]
    h = engine.prepare(s, th)
    res = engine.count_prepared([h], th, variance=True)[0]
    pr = prep.prepare(d)
    jp, js = prep.jackknife_region_indices(pr["positions"], pr["shapes"], L, 3)
    ref = brute.count_pairs(pr, r_bins, sb, R_MIN, R_MAX, L, jk_pos=jp, jk_shape=js, num_box=27)
    compare(res, ref, p.R, 27, weighted=True)


@pytest.mark.parametrize("mode", ["rppi", "rmu"])
def test_graph_replay_equals_plain_launches(engine, mode):
    import torch
    th, _ = _th(mode)
    sm = _samples()
    dev = torch.device("cuda", 0)

    def on_device(s):
        t = lambda a: None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
        return SampleArrays(positions=t(s.positions), shapes=t(s.shapes), axis=t(s.axis), e=t(s.e),
                            weight=t(s.weight), weight_shape=t(s.weight_shape), boxsize=s.boxsize,
                            los=s.los, jk_per_side=s.jk_per_side, jk_sub=s.jk_sub)

    eager = engine.count_batch([s for _, _, s in sm], th, tuning={"graph": -1})
    assert all(r.stats["graph"] == 0 for r in eager)
    replays = 0
    for k in range(4):
        # fresh device copies each time: the replay must take the new pointers
        batch = [on_device(s) for _, _, s in sm]
        got = engine.count_batch(batch, th)
        replays += got[0].stats["graph"]
        for i, (g, e) in enumerate(zip(got, eager)):
            _same(g, e, f"replay {k} sample {i}")
        # eager work and a sync between replays
        engine.count(sm[0][2], th, method="fast")
        torch.cuda.synchronize()
    assert replays >= 3, replays
