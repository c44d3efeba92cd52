"""MeasureIABox end to end on the GPU against the oracle-built expected file.

Every dataset the reference would write must be present (and nothing else),
with gg quantities (DD, RR, xi_gg, w_gg, their realisations and covariance)
bit-exact for unit weights and every g+ / gx quantity within 1e-6.
"""
import numpy as np
import pytest

from measure_ia_amd import MeasureIABox, ReadData
from measure_ia_amd.hdf5io import H5Store
from measure_ia_amd.synth import ellipticity_noise, shapes_from_components, wrap_box
from refbox import expected_output

pytestmark = pytest.mark.gpu

L = 60.0


def _data(seed=0, n_pos=2500, n_shape=1200, weights=False, los=2):
    rng = np.random.default_rng(seed)
    pos = wrap_box(rng.uniform(0, L, (n_pos, 3)), L)
    shp = wrap_box(rng.uniform(0, L, (n_shape, 3)), L)
    axis, q = shapes_from_components(ellipticity_noise(rng, n_shape, 0.25))
    d = {"Position": pos, "Position_shape_sample": shp, "Axis_Direction": axis, "LOS": los, "q": q}
    if weights:
        d["weight"] = rng.uniform(0.5, 1.5, n_pos)
        d["weight_shape_sample"] = rng.uniform(0.5, 1.5, n_shape)
    return d


def _read_all(fn):
    with H5Store(fn, "r") as f:
        return {p: f.read(p) for p in f.walk("")}


def _compare(got, exp, exact_gg=True):
    assert set(got) == set(exp), (sorted(set(got) ^ set(exp)))[:10]
    for k, ref in exp.items():
        g = got[k]
        assert g.shape == ref.shape, k
        if exact_gg and not any(s in k for s in ("g_plus", "g_cross", "sigmasq")):
            np.testing.assert_array_equal(g, ref, err_msg=k)
        else:
            scale = np.abs(ref).max() if ref.size else 0.0
            np.testing.assert_allclose(g, ref, rtol=1e-6, atol=1e-6 * scale + 1e-300, err_msg=k)


CASES = [
    ("w", "both", 8, "brute", None),
    ("w", "both", 8, "tree", None),
    ("w", "both", 27, "multiprocessing", None),
    ("w", "g+", 0, "tree", None),
    ("w", "gg", 0, "brute", None),
    ("m", "both", 8, "tree", None),
    ("m", "both", 27, "brute", None),
    ("m", "gg", 0, "tree", None),
]


@pytest.mark.parametrize("grp, corr, num_jk, layout, _", CASES)
def test_box_matches_reference_layout(tmp_path, grp, corr, num_jk, layout, _):
    d = _data(seed=num_jk + len(corr))
    fn = str(tmp_path / "out.hdf5")
    box = MeasureIABox(d, fn, None, 99, [0.1, 10.0], 10, 8, 20.0, boxsize=L,
                       num_nodes=4 if layout == "multiprocessing" else 1, verbose=False)
    tfp = False if layout == "brute" else str(tmp_path)
    if grp == "w":
        box.measure_xi_w("All", corr, num_jk, temp_file_path=tfp)
    else:
        box.measure_xi_multipoles("All", corr, num_jk, temp_file_path=tfp)
    exp = expected_output(d, L, "Snapshot_99/", [0.1, 10.0], 10, 8, 20.0, grp, "All", corr, num_jk,
                          layout)
    _compare(_read_all(fn), exp)
    # the reader finds what the writer wrote
    rd = ReadData("TNG300", "out", 99, data_path=str(tmp_path))
    rd.read_MeasureIA_output("All", num_jk if num_jk else None)
    if grp == "w" and corr in ("both", "gg"):
        assert rd.w_gg is not None and rd.rp is not None
    if grp == "m" and corr in ("both", "g+"):
        assert rd.multipoles_gp is not None


def test_masks_weights_los(tmp_path):
    d = _data(seed=5, weights=True, los=0)
    rng = np.random.default_rng(1)
    m = {"Position": rng.uniform(size=2500) < 0.7, "Position_shape_sample": rng.uniform(size=1200) < 0.6}
    m["Axis_Direction"] = m["Position_shape_sample"]
    m["q"] = m["Position_shape_sample"]
    m["weight"] = m["Position"]
    m["weight_shape_sample"] = m["Position_shape_sample"]
    fn = str(tmp_path / "out.hdf5")
    box = MeasureIABox(d, fn, None, None, [0.1, 10.0], 10, 8, 20.0, boxsize=L, verbose=False)
    box.measure_xi_w("sel", "both", 8, temp_file_path=None, masks=dict(m), ellipticity="ellipticity")
    exp = expected_output(d, L, "", [0.1, 10.0], 10, 8, 20.0, "w", "sel", "both", 8, "tree",
                          masks=dict(m), ellipticity="ellipticity")
    _compare(_read_all(fn), exp, exact_gg=False)


@pytest.mark.parametrize("grp", ["w", "m"])
def test_open_box_matches_reference_layout(tmp_path, grp):
    """MeasureIABox(periodicity=False): separations are not wrapped."""
    d = _data(seed=13)
    fn = str(tmp_path / "out.hdf5")
    box = MeasureIABox(d, fn, None, None, [0.1, 10.0], 10, 8, 20.0, boxsize=L, periodicity=False,
                       verbose=False)
    if grp == "w":
        box.measure_xi_w("open", "both", 8, temp_file_path=str(tmp_path))
    else:
        box.measure_xi_multipoles("open", "both", 8, temp_file_path=str(tmp_path))
    exp = expected_output(d, L, "", [0.1, 10.0], 10, 8, 20.0, grp, "open", "both", 8, "tree",
                          periodic=False)
    _compare(_read_all(fn), exp)


def test_weight_scaling_like_reference(tmp_path):
    """tests/test_weights.py:5-43 of the reference: weights 0.5 -> w and DD / 4, cov / 16."""
    d = _data(seed=7)
    fn = str(tmp_path / "out.hdf5")
    box = MeasureIABox(d, fn, None, 99, [0.1, 10.0], 10, 8, None, boxsize=L, verbose=False)
    box.measure_xi_w("All_both", "both", 8, temp_file_path=str(tmp_path))
    box.data["weight"] = np.array([0.5] * len(d["Position"]))
    box.data["weight_shape_sample"] = np.array([0.5] * len(d["Position_shape_sample"]))
    box.measure_xi_w("All_both_weight", "both", 8, temp_file_path=str(tmp_path))
    out = _read_all(fn)
    g = "Snapshot_99"
    # DD is exact.  The e+ sums are FP64 additions in an unordered (atomic)
    # sequence, so the reference's bit-exact x4 holds to FP64 rounding only.
    np.testing.assert_array_equal(out[f"{g}/w/xi_gg/All_both_DD"], 4 * out[f"{g}/w/xi_gg/All_both_weight_DD"])
    np.testing.assert_allclose(out[f"{g}/w_g_plus/All_both"], 4 * out[f"{g}/w_g_plus/All_both_weight"],
                               rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(out[f"{g}/w_g_plus/All_both_jackknife_cov_8"],
                               16 * out[f"{g}/w_g_plus/All_both_weight_jackknife_cov_8"])
    np.testing.assert_allclose(out[f"{g}/w_gg/All_both_jackknife_cov_8"],
                               16 * out[f"{g}/w_gg/All_both_weight_jackknife_cov_8"])


def test_corr_types_agree(tmp_path):
    """tests/test_w_sim_internal_consistency.py:35-56: g+ / gg alone equal 'both'."""
    d = _data(seed=9)
    fn = str(tmp_path / "out.hdf5")
    box = MeasureIABox(d, fn, None, 99, [0.1, 10.0], 10, 8, None, boxsize=L, verbose=False)
    for name, corr in (("All_both", "both"), ("All_gp", "g+"), ("All_gg", "gg")):
        box.measure_xi_w(name, corr, 0, temp_file_path=str(tmp_path))
    out = _read_all(fn)
    g = "Snapshot_99"
    # gg is exact; g+ comes from two separate counts whose FP64 atomic sums
    # may be added in different orders
    np.testing.assert_allclose(out[f"{g}/w_g_plus/All_both"], out[f"{g}/w_g_plus/All_gp"], rtol=1e-12,
                               atol=1e-15)
    np.testing.assert_array_equal(out[f"{g}/w_gg/All_both"], out[f"{g}/w_gg/All_gg"])


def test_plain_and_fast_backends_agree(tmp_path):
    d = _data(seed=11)
    outs = []
    for be in ("hip", "hip-plain"):
        fn = str(tmp_path / f"{be}.hdf5")
        box = MeasureIABox(d, fn, None, None, [0.1, 10.0], 10, 8, 20.0, boxsize=L, backend=be,
                           verbose=False)
        box.measure_xi_w("All", "both", 27, temp_file_path=False)
        outs.append(_read_all(fn))
    _compare(outs[0], outs[1])


def test_repeated_calls_reuse_the_device_sort(tmp_path):
    """A second call on unchanged data re-uses the host prep, device copy and
    device sort (mia_prepare_sample); a change to the data, even in place,
    is seen and counted afresh."""
    d = _data(seed=17)
    fn = str(tmp_path / "out.hdf5")
    box = MeasureIABox(d, fn, None, None, [0.1, 10.0], 10, 8, 20.0, boxsize=L, verbose=False)
    box.measure_xi_w("a", "both", 8, temp_file_path=str(tmp_path))
    assert box.last_timing["engine"]["reused_sort"] is False
    box.measure_xi_multipoles("m", "both", 8, temp_file_path=str(tmp_path))
    assert box.last_timing["engine"]["reused_sort"] is False  # another binning: its own sort
    box.measure_xi_w("b", "both", 8, temp_file_path=str(tmp_path))
    assert box.last_timing["engine"]["reused_sort"] is True
    out = _read_all(fn)
    exp_w = expected_output(d, L, "", [0.1, 10.0], 10, 8, 20.0, "w", "b", "both", 8, "tree")
    exp_m = expected_output(d, L, "", [0.1, 10.0], 10, 8, 20.0, "m", "m", "both", 8, "tree")
    _compare({k: v for k, v in out.items() if k in exp_w}, exp_w)
    _compare({k: v for k, v in out.items() if k in exp_m}, exp_m)
    # in-place change of one position: a fresh sort of the new data
    box.data["Position"][5] = (box.data["Position"][5] + 7.0) % L
    fn2 = str(tmp_path / "out2.hdf5")
    box.output_file_name = fn2
    box.measure_xi_w("c", "both", 8, temp_file_path=str(tmp_path))
    assert box.last_timing["engine"]["reused_sort"] is False
    exp_c = expected_output(box.data, L, "", [0.1, 10.0], 10, 8, 20.0, "w", "c", "both", 8, "tree")
    _compare(_read_all(fn2), exp_c)
