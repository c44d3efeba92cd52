"""Error of every output quantity against the FP64 oracle, stated per quantity.

One clustered, weighted box (4,000 positions / 1,600 shapes, 27 regions) is
measured through MeasureIABox for w_gg / w_g+ and for the multipoles, and
every dataset is compared with the oracle-built expected file.  For each
family of datasets the test prints

  max |got - ref| / max|ref|      (absolute error in units of the largest value)
  max |got - ref| / |ref|         (relative error over bins with |ref| > 1e-3 max|ref|)

and holds them to the bar: pair counts equal (unit weights) and every other
quantity within 1e-6.  Run with ``pytest -s`` to see the table.
"""
import numpy as np
import pytest

from measure_ia_amd import MeasureIABox
from measure_ia_amd.hdf5io import H5Store
from measure_ia_amd.synth import clustered
from refbox import expected_output

pytestmark = pytest.mark.gpu

L = 90.0


def _families(grp):
    if grp == "w":
        return {
            "DD": ["/xi_gg/", "_DD"], "S+D": ["_SplusD"], "SxD": ["_ScrossD"],
            "xi_gg": ["w/xi_gg/"], "xi_g+": ["w/xi_g_plus/"], "xi_gx": ["w/xi_g_cross/"],
            "w_gg": ["w_gg/"], "w_g+": ["w_g_plus/"], "cov w_gg": ["w_gg/", "jackknife_cov"],
            "cov w_g+": ["w_g_plus/", "jackknife_cov"],
        }
    return {
        "DD": ["/xi_gg/", "_DD"], "S+D": ["_SplusD"], "xi_gg(r,mu)": ["multipoles/xi_gg/"],
        "xi_g+(r,mu)": ["multipoles/xi_g_plus/"], "multipole gg l=0": ["multipoles_gg/"],
        "multipole g+ l=2": ["multipoles_g_plus/"], "cov multipole g+": ["multipoles_g_plus/", "jackknife_cov"],
    }


def _report(got, exp, grp, title):
    rows = []
    for fam, keys in _families(grp).items():
        paths = [p for p in exp if all(k in p for k in keys) and not p.endswith(("_rp", "_pi", "_r", "_mu_r"))
                 and "sigmasq" not in p and "_RR" not in p]
        if fam.startswith("xi"):
            paths = [p for p in paths if not any(s in p.split("/")[-1] for s in ("_DD", "_SplusD", "_ScrossD"))]
        assert paths, fam
        ea = er = 0.0
        for p in paths:
            g, r = np.asarray(got[p], float), np.asarray(exp[p], float)
            m = np.abs(r).max()
            if m == 0:
                continue
            d = np.abs(g - r)
            ea = max(ea, d.max() / m)
            big = np.abs(r) > 1e-3 * m
            if big.any():
                er = max(er, (d[big] / np.abs(r[big])).max())
        rows.append((fam, len(paths), ea, er))
    print(f"\n{title}")
    print(f"{'quantity':22s} {'datasets':>8s} {'max|d|/max|ref|':>16s} {'max rel':>10s}")
    for fam, n, ea, er in rows:
        print(f"{fam:22s} {n:8d} {ea:16.2e} {er:10.2e}")
    return {fam: (ea, er) for fam, n, ea, er in rows}


def _read(fn):
    with H5Store(fn, "r") as f:
        return {p: f.read(p) for p in f.walk("")}


@pytest.mark.parametrize("grp", ["w", "m"])
def test_error_of_every_quantity(tmp_path, grp):
    cat = clustered(4000, 1600, L, 60, 1.0, 40, 0.2, 0.2, seed=5)
    d = cat.data
    fn = str(tmp_path / "out.hdf5")
    box = MeasureIABox(d, fn, None, 99, [0.1, 15.0], 8, 10, 30.0, boxsize=L, verbose=False)
    if grp == "w":
        box.measure_xi_w("All", "both", 27, temp_file_path=str(tmp_path))
    else:
        box.measure_xi_multipoles("All", "both", 27, temp_file_path=str(tmp_path))
    exp = expected_output(d, L, "Snapshot_99/", [0.1, 15.0], 8, 10, 30.0, grp, "All", "both", 27, "tree")
    got = _read(fn)
    errs = _report(got, exp, grp, f"errors against the FP64 oracle ({'w' if grp == 'w' else 'multipoles'})")
    assert errs["DD"] == (0.0, 0.0)
    for fam, (ea, er) in errs.items():
        assert ea <= 1e-6, fam
