"""Entry points without a GPU.

* bench.py asked for more than one GPU exits with status 2 and says why (the
  counter runs on one GPU; sharding is described, not built);
* __graft_entry__ exposes build() and smoke();
* the product has no CPU path: with no HIP device visible, a box and a
  lightcone measurement raise instead of counting on the host.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_bench_lists_the_workloads():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0
    for wl in ("clustered1M", "uniform20k", "multipoles2M", "batch12", "large20M", "lightcone1M",
               "lightcone_rr4M"):
        assert wl in r.stdout


def test_graft_entry_points():
    sys.path.insert(0, ROOT)
    import __graft_entry__ as g
    assert callable(g.build) and callable(g.smoke)


def _no_gpu():
    import torch
    return not torch.cuda.is_available()


@pytest.mark.skipif(not _no_gpu(), reason="a HIP device is visible")
def test_box_measurement_has_no_cpu_path(tmp_path):
    from measure_ia_amd import MeasureIABox
    from measure_ia_amd.synth import uniform
    cat = uniform(500, 50.0, seed=0)
    box = MeasureIABox(cat.data, str(tmp_path / "x.hdf5"), None, None, [0.5, 5.0], 4, 4, 10.0,
                       boxsize=cat.boxsize, verbose=False)
    with pytest.raises(RuntimeError, match="no HIP device"):
        box.measure_xi_w("d", "both", 8, temp_file_path=None)


@pytest.mark.skipif(not _no_gpu(), reason="a HIP device is visible")
def test_lightcone_measurement_has_no_cpu_path(tmp_path):
    from measure_ia_amd import MeasureIALightcone
    rng = np.random.default_rng(0)
    n = 200
    data = {"RA": rng.uniform(10, 12, n), "DEC": rng.uniform(-1, 1, n), "Redshift": rng.uniform(0.1, 0.2, n),
            "RA_shape_sample": rng.uniform(10, 12, n), "DEC_shape_sample": rng.uniform(-1, 1, n),
            "Redshift_shape_sample": rng.uniform(0.1, 0.2, n), "e1": rng.normal(0, 0.2, n),
            "e2": rng.normal(0, 0.2, n), "weight": np.ones(n), "weight_shape_sample": np.ones(n)}
    randoms = {"RA": rng.uniform(10, 12, n), "DEC": rng.uniform(-1, 1, n), "Redshift": rng.uniform(0.1, 0.2, n)}
    lc = MeasureIALightcone(data, randoms, str(tmp_path / "x.hdf5"), separation_limits=[0.5, 10.0],
                            num_bins_r=4, num_bins_pi=4, pi_max=10.0)
    jk = {"position": np.zeros(n, int), "shape": np.zeros(n, int), "randoms": np.zeros(n, int)}
    with pytest.raises(RuntimeError, match="no HIP device"):
        lc.measure_xi_w("galaxies", "gp", "both", jk_patches=jk)
