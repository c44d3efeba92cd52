"""Reading MeasureIA output files (the output half of ReadData, read_data.py:41-118, 304-374).

``ReadData(simulation, catalogue, snapshot, sub_group, data_path=...)``
reads ``{data_path}/{catalogue}.hdf5``; ``read_cat`` returns one dataset
under ``Snapshot_{n}/{sub_group}/`` and ``read_MeasureIA_output`` fills the
w / multipole / covariance attributes of one dataset name.  The raw
simulation readers of the reference (subhalo and snapshot ingestion) are not
part of this package.
"""
from __future__ import annotations

import os

from .hdf5io import H5Store


class ReadData:
    _ATTRS = ("r", "rp", "w_gg", "w_gp", "multipoles_gg", "multipoles_gp", "cov_multipoles_gg",
              "errors_multipoles_gg", "cov_multipoles_gp", "errors_multipoles_gp", "cov_w_gg",
              "errors_w_gg", "cov_w_gp", "errors_w_gp")

    def __init__(self, simulation, catalogue, snapshot, sub_group="", output_file_name=None,
                 data_path="./data/raw/"):
        self.simname = simulation
        self.snapshot = None if snapshot is None else str(snapshot)
        self.snap_group = "" if snapshot is None else f"Snapshot_{snapshot}/"
        self.catalogue = catalogue
        self.sub_group = sub_group + "/"
        self.data_path = data_path + "/"
        self.output_file_name = output_file_name
        for a in self._ATTRS:
            setattr(self, a, None)

    def _file(self):
        return os.path.join(self.data_path, f"{self.catalogue}.hdf5")

    def read_cat(self, dataset_name, cut=None):
        """One dataset (or the slice [cut[0]:cut[1]]) of the catalogue file."""
        if self.catalogue == "Subhalo":
            raise KeyError("Use read_subhalo method")
        if self.catalogue == "Snapshot":
            raise KeyError("Use read_snapshot method")
        with H5Store(self._file(), "r") as f:
            data = f.read(self.snap_group + self.sub_group + dataset_name)
        return data if cut is None else data[cut[0]:cut[1]]

    def read_MeasureIA_output(self, dataset_name, num_jk):
        """Fill w_gg, w_gp, multipoles_gg/gp, r, rp and their jackknife cov/errors."""
        for a in self._ATTRS:
            setattr(self, a, None)
        with H5Store(self._file(), "r") as f:
            base = self.snap_group

            def get(path):
                return f.read(base + path)

            # assignments happen one by one inside each try, as in the reference:
            # a missing covariance leaves the values already read in place
            for grp, attr, sep, sep_attr in (("multipoles_gg", "multipoles_gg", "_r", "r"),
                                             ("multipoles_g_plus", "multipoles_gp", "_r", "r"),
                                             ("w_gg", "w_gg", "_rp", "rp"),
                                             ("w_g_plus", "w_gp", "_rp", "rp")):
                try:
                    setattr(self, attr, get(f"{grp}/{dataset_name}"))
                    setattr(self, sep_attr, get(f"{grp}/{dataset_name}{sep}"))
                    if num_jk is not None:
                        setattr(self, f"cov_{attr}", get(f"{grp}/{dataset_name}_jackknife_cov_{num_jk}"))
                        setattr(self, f"errors_{attr}", get(f"{grp}/{dataset_name}_jackknife_{num_jk}"))
                except KeyError:
                    pass
