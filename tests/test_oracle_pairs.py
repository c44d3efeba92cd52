"""Pin the oracle's pair counters (oracle/brute.py, oracle/tree.py) without the reference.

The raw catalogue behind the reference's pair-level tests is not shipped, so
the counters are pinned by checks a plausible slip would fail:
  * an independent per-pair Python loop (math.* scalars) on tiny inputs;
  * hand-placed pairs with known (r_p, pi) and axes at 0, 45, 90 degrees,
    where e+ = +e, 0, -e (the box convention, measure_IA_base.py:226);
  * uniform randoms: DD per bin equals the analytic RR within Poisson error;
  * weights 0.5 -> DD and S+D divided by 4 (tests/test_weights.py:5-84);
  * jackknife: realisation i equals a full count with region i removed;
  * the KDTree counter equals the brute counter.
"""
import numpy as np
import pytest

from oracle import brute, post, prep, tree


def _sample(n_pos, n_shape, L, seed, los=2, weights=False):
    rng = np.random.default_rng(seed)
    d = {"Position": rng.uniform(0, L, (n_pos, 3)), "Position_shape_sample": rng.uniform(0, L, (n_shape, 3)),
         "Axis_Direction": rng.normal(size=(n_shape, 2)), "LOS": los,
         "q": rng.uniform(0.3, 1.0, n_shape)}
    if weights:
        d["weight"] = rng.uniform(0.5, 1.5, n_pos)
        d["weight_shape_sample"] = rng.uniform(0.5, 1.5, n_shape)
    return d


@pytest.mark.parametrize("mode", ["rppi", "rmu"])
@pytest.mark.parametrize("los", [0, 1, 2])
def test_brute_equals_scalar_loop(mode, los):
    L = 10.0
    d = _sample(60, 40, L, seed=los, los=los, weights=True)
    pr = prep.prepare(d)
    r_bins, pi_bins, mu_bins = prep.bin_edges(0.2, 4.0, 5, 6, 3.0, L)
    sb = pi_bins if mode == "rppi" else mu_bins
    ref = brute.count_pairs(pr, r_bins, sb, 0.2, 4.0, L, mode=mode)
    sc = brute.count_pairs_scalar(pr, r_bins, sb, 0.2, 4.0, L, mode=mode)
    R = pr["R"]
    got = np.zeros_like(ref["DD"])
    gp = np.zeros_like(ref["DD"])
    gx = np.zeros_like(ref["DD"])
    for (i, j), (dd, ep, ex) in sc.items():
        got[i, j], gp[i, j], gx[i, j] = dd, ep, ex
    assert ref["npairs"] > 50
    np.testing.assert_allclose(ref["DD"], got, rtol=1e-13)
    np.testing.assert_allclose(ref["SplusD"], gp / (2 * R), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(ref["ScrossD"], gx / (2 * R), rtol=1e-9, atol=1e-12)


def _known_case():
    """One shape at the box centre, positions at chosen separations."""
    L = 20.0
    c = np.array([10.0, 10.0, 10.0])
    # (dx, dy, dz) from position to shape: shape - position = sep
    seps = np.array([[1.5, 0.0, 0.5],     # r_p 1.5 along x
                     [0.0, 3.0, -2.5],    # r_p 3 along y
                     [2.0, 2.0, 1.0],     # r_p 2.83 at 45 deg
                     [0.05, 0.0, 0.0],    # r_p below r_min: excluded
                     [9.0, 0.0, 0.0]])    # beyond r_max: excluded
    pos = c - seps
    return L, c, pos


@pytest.mark.parametrize("angle, expect", [(0.0, [1, -1, 0]), (np.pi / 2, [-1, 1, 0]),
                                           (np.pi / 4, [0, 0, 1])])
def test_hand_placed_known_answers(angle, expect):
    L, c, pos = _known_case()
    q = 0.5
    e = (1 - q ** 2) / (1 + q ** 2)
    d = {"Position": pos, "Position_shape_sample": c[None, :], "LOS": 2, "q": np.array([q]),
         "Axis_Direction": np.array([[np.cos(angle), np.sin(angle)]])}
    pr = prep.prepare(d)
    r_bins, pi_bins, _ = prep.bin_edges(0.1, 5.0, 4, 4, 4.0, L)
    out = brute.count_pairs(pr, r_bins, pi_bins, 0.1, 5.0, L)
    assert out["npairs"] == 3
    # bins: r edges 0.1, 0.266, 0.707, 1.88, 5 ; pi edges -4, -2, 0, 2, 4
    exp_bins = [(2, 2), (3, 0), (3, 2)]
    for (ir, ip), sgn in zip(exp_bins, expect):
        assert out["DD"][ir, ip] == 1.0
        assert out["SplusD"][ir, ip] * 2 * pr["R"] == pytest.approx(sgn * e, abs=1e-12)
    assert out["DD"].sum() == 3
    # a planted sign slip (the lightcone convention -e cos 2phi) fails this check
    if expect[0] != 0:
        assert not np.isclose(-out["SplusD"][2, 2] * 2 * pr["R"], expect[0] * e)


def test_uniform_dd_matches_analytic_rr():
    L = 40.0
    rng = np.random.default_rng(7)
    n = 3000
    d = {"Position": rng.uniform(0, L, (n, 3)), "Position_shape_sample": rng.uniform(0, L, (n, 3)),
         "Axis_Direction": rng.normal(size=(n, 2)), "LOS": 2, "q": rng.uniform(0.3, 1, n)}
    pr = prep.prepare(d)
    r_bins, pi_bins, mu_bins = prep.bin_edges(0.5, 8.0, 4, 4, 8.0, L)
    out = brute.count_pairs(pr, r_bins, pi_bins, 0.5, 8.0, L)
    rr = post.rr_grid("rppi", r_bins, pi_bins, L ** 3, n, n)
    z = (out["DD"] - rr) / np.sqrt(rr)
    assert np.abs(z).max() < 5, z
    # (r, mu): uniform pairs are uniform in mu; the expectation is the spherical
    # shell times dmu/2.  (The reference's own RR uses spherical-cap volumes,
    # measure_IA_base.py:274-351, which is not uniform in mu; the fixtures pin
    # that form, see test_oracle_post.py.)
    out = brute.count_pairs(pr, r_bins, mu_bins, 0.5, 8.0, L, mode="rmu")
    shell = 4 * np.pi / 3 * (r_bins[1:] ** 3 - r_bins[:-1] ** 3)
    expect = n * n * shell[:, None] * (np.diff(mu_bins) / 2)[None, :] / L ** 3
    z = (out["DD"] - expect) / np.sqrt(expect)
    assert np.abs(z).max() < 5, z


def test_weight_scaling():
    L = 15.0
    d = _sample(300, 200, L, seed=2)
    r_bins, pi_bins, _ = prep.bin_edges(0.2, 5.0, 5, 4, 5.0, L)
    a = brute.count_pairs(prep.prepare(d), r_bins, pi_bins, 0.2, 5.0, L)
    d["weight"] = np.full(300, 0.5)
    d["weight_shape_sample"] = np.full(200, 0.5)
    b = brute.count_pairs(prep.prepare(d), r_bins, pi_bins, 0.2, 5.0, L)
    np.testing.assert_array_equal(a["DD"], 4 * b["DD"])
    np.testing.assert_allclose(a["SplusD"], 4 * b["SplusD"], rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("mode", ["rppi", "rmu"])
def test_jackknife_leave_one_out_identity(mode):
    """DD - DD_jk[i] equals DD of the sample with region i removed (both samples)."""
    L = 12.0
    d = _sample(700, 400, L, seed=5)
    d["Position"][:40] = np.round(d["Position"][:40] / 6.0) * 6.0  # objects on faces -> region 0
    pr = prep.prepare(d)
    r_bins, pi_bins, mu_bins = prep.bin_edges(0.2, 5.0, 5, 4, 5.0, L)
    sb = pi_bins if mode == "rppi" else mu_bins
    jp, js = prep.jackknife_region_indices(pr["positions"], pr["shapes"], L, 2)
    full = brute.count_pairs(pr, r_bins, sb, 0.2, 5.0, L, mode=mode, jk_pos=jp, jk_shape=js, num_box=8)
    for i in range(8):
        sub = dict(pr)
        mp, ms = jp != i, js != i
        for k in ("positions", "weight"):
            sub[k] = pr[k][mp]
        for k in ("shapes", "axis", "e", "weight_shape"):
            sub[k] = pr[k][ms]
        part = brute.count_pairs(sub, r_bins, sb, 0.2, 5.0, L, mode=mode)
        np.testing.assert_array_equal(full["DD"] - full["DD_jk"][i], part["DD"])
        np.testing.assert_allclose(full["SplusD"] * 2 * pr["R"] - full["SplusD_jk"][i],
                                   part["SplusD"] * 2 * pr["R"], rtol=1e-9, atol=1e-11)
        # a planted slip -- tallying a pair twice when shape and position share
        # the region -- adds DD(region i alone) to the tally and breaks the identity
        if i == 0:
            only = dict(pr)
            for k in ("positions", "weight"):
                only[k] = pr[k][~mp]
            for k in ("shapes", "axis", "e", "weight_shape"):
                only[k] = pr[k][~ms]
            same = brute.count_pairs(only, r_bins, sb, 0.2, 5.0, L, mode=mode)["DD"]
            assert same.sum() > 0
            assert not np.array_equal(full["DD"] - (full["DD_jk"][i] + same), part["DD"])


@pytest.mark.parametrize("mode", ["rppi", "rmu"])
def test_tree_equals_brute(mode):
    L = 30.0
    d = _sample(1500, 600, L, seed=9, weights=True)
    pr = prep.prepare(d)
    r_bins, pi_bins, mu_bins = prep.bin_edges(0.1, 6.0, 6, 6, 6.0, L)
    sb = pi_bins if mode == "rppi" else mu_bins
    jp, js = prep.jackknife_region_indices(pr["positions"], pr["shapes"], L, 3)
    a = brute.count_pairs(pr, r_bins, sb, 0.1, 6.0, L, mode=mode, jk_pos=jp, jk_shape=js, num_box=27)
    b = tree.count_pairs(pr, r_bins, sb, 0.1, 6.0, L, mode=mode, jk_pos=jp, jk_shape=js, num_box=27,
                         chunk_size=100, num_nodes=2)
    np.testing.assert_allclose(a["DD"], b["DD"], rtol=1e-13)
    np.testing.assert_allclose(a["DD_jk"], b["DD_jk"], rtol=1e-13)
    np.testing.assert_allclose(a["SplusD"], b["SplusD"], rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(a["SplusD_jk"], b["SplusD_jk"], rtol=1e-10, atol=1e-13)
    assert a["npairs"] == b["npairs"]


@pytest.mark.parametrize("mode", ["rppi", "rmu"])
def test_open_box_scalar_loop_and_subset_of_periodic(mode):
    """periodicity=False (measure_w_box_jk.py:160-162: no wrap).

    The brute counter equals the scalar loop without wrap, and every pair it
    keeps is also kept, in the same bin, by the periodic count (an unwrapped
    in-domain separation is below L/2, where the wrap changes nothing), while
    the periodic count finds more pairs across the faces.
    """
    L = 10.0
    d = _sample(80, 50, L, seed=7, weights=True)
    pr = prep.prepare(d)
    r_bins, pi_bins, mu_bins = prep.bin_edges(0.2, 4.0, 5, 6, 3.0, L)
    sb = pi_bins if mode == "rppi" else mu_bins
    op = brute.count_pairs(pr, r_bins, sb, 0.2, 4.0, L, mode=mode, periodic=False)
    sc = brute.count_pairs_scalar(pr, r_bins, sb, 0.2, 4.0, L, mode=mode, periodic=False)
    got = np.zeros_like(op["DD"])
    for (i, j), (dd, ep, ex) in sc.items():
        got[i, j] = dd
    np.testing.assert_allclose(op["DD"], got, rtol=1e-13)
    per = brute.count_pairs(pr, r_bins, sb, 0.2, 4.0, L, mode=mode, periodic=True)
    assert np.all(op["NP"] <= per["NP"])
    assert op["NP"].sum() < per["NP"].sum()


@pytest.mark.parametrize("n", [1, 2, 3, 5])
def test_region_ids_by_axis_equal_the_triple_loop(n):
    """The per-axis form of the region test gives the literal loop's ids,
    also for objects on a face, at 0 and at L, and one ulp either side of a face."""
    rng = np.random.default_rng(40 + n)
    L = 205.0
    x = rng.uniform(0, L, (6000, 3))
    face = np.arange(n + 1) * (L / 2.0 * 2.0 / n)
    k = rng.integers(0, n + 1, (1500, 3))
    x[:1500] = face[k]
    x[1500:2000] = np.nextafter(face[k[:500]], 0.0)
    x[2000:2500] = np.nextafter(face[k[500:1000]], L + 1)
    x[2500:2600] = 0.0
    x[2600:2700] = L
    x = np.clip(x, 0.0, L)
    rng.shuffle(x, axis=1)
    a_p, a_s = prep.jackknife_region_indices(x, x[::7], L, n)
    b_p, b_s = prep.jackknife_region_indices_by_axis(x, x[::7], L, n)
    np.testing.assert_array_equal(b_p, a_p)
    np.testing.assert_array_equal(b_s, a_s)
    assert (a_p > 0).sum() > 1000 or n == 1


def test_parallel_axes_case_pins_the_arccos_nan_rule():
    """The parallel_axes case makes the reference's d round outside [-1, 1].

    measure_w_box_jk.py:166-175: d = 0 + (dx/r_p) a_x, then + (dy/r_p) a_y;
    arccos(d) is NaN for |d| > 1 and isnan -> e+ = ex = 0.  Recomputed here
    pair by pair with an explicit mask and the closed form e (2 d^2 - 1), the
    oracle's S+D must agree; dropping the rule (using the closed form for
    every pair) must not.
    """
    from paircases import make_case
    data, L, cfg = make_case("parallel_axes")
    pr = prep.prepare(data)
    r_bins, pi_bins, _ = prep.bin_edges(cfg["r_min"], cfg["r_max"], cfg["nr"], cfg["n2"], cfg["pi_max"], L)
    out = brute.count_pairs(pr, r_bins, pi_bins, cfg["r_min"], cfg["r_max"], L)
    sep = pr["shapes"][None, :, :] - pr["positions"][:, None, :]
    sep[sep > L / 2] -= L
    sep[sep < -L / 2] += L
    a, b, z = sep[..., pr["not_los"][0]], sep[..., pr["not_los"][1]], sep[..., pr["los"]]
    rp = np.sqrt(a * a + b * b)
    m = (rp >= r_bins[0]) & (rp < r_bins[-1]) & (z >= pi_bins[0]) & (z < pi_bins[-1])
    with np.errstate(invalid="ignore", divide="ignore"):
        d = 0.0 + (a / rp) * pr["axis"][None, :, 0]
        d = d + (b / rp) * pr["axis"][None, :, 1]
    nan_pairs = m & (np.abs(d) > 1.0)
    assert nan_pairs.sum() >= 50, nan_pairs.sum()
    e = np.broadcast_to(pr["e"][None, :], d.shape)
    closed = e * (2 * d * d - 1)
    keep = m & ~nan_pairs
    with_rule = closed[keep].sum() / (2 * pr["R"])
    without = closed[m].sum() / (2 * pr["R"])
    got = out["SplusD"].sum()
    scale = np.abs(closed[m]).sum() / (2 * pr["R"])
    assert abs(got - with_rule) <= 1e-12 * scale
    assert abs(got - without) > 1e-3 * scale
