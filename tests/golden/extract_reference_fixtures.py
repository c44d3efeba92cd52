"""Extract the reference's shipped output fixtures into compressed .npz files.

Source data (read-only, data files held by the reference's own tests):
    tests/data/processed/TNG300/mock_IA_TNG300.hdf5        (N=766 "All")
    tests/data/processed/TNG300/mock_IA_TNG300_large.hdf5  (N=4547 "All")
They are the expected outputs of the reference's test suite
(tests/test_w_sim_internal_consistency.py:19-30, tests/test_w_jk.py:7-12).
Every dataset is copied verbatim (float64); keys are the HDF5 paths with
'/' replaced by '|'.  The reference tree is only needed to regenerate the
fixtures; tests read the committed .npz files.

Run:  python tests/golden/extract_reference_fixtures.py [/root/reference]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from measure_ia_amd.hdf5io import H5Store  # noqa: E402

FILES = {
    "tng300_mock_small.npz": "tests/data/processed/TNG300/mock_IA_TNG300.hdf5",
    "tng300_mock_large.npz": "tests/data/processed/TNG300/mock_IA_TNG300_large.hdf5",
}


def main(ref_root="/root/reference"):
    for out_name, rel in FILES.items():
        src = os.path.join(ref_root, rel)
        arrays = {}
        with H5Store(src, "r") as f:
            for path in f.walk(""):
                arrays[path.replace("/", "|")] = f.read(path)
        np.savez_compressed(os.path.join(HERE, out_name), **arrays)
        print(f"{out_name}: {len(arrays)} datasets")


if __name__ == "__main__":
    main(*sys.argv[1:])
