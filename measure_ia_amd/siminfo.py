"""Simulation table: box size, L/2, h and the snapshot group (Sim_info.py:37-135)."""
from __future__ import annotations


_BOXES = {
    # name: (boxsize in cMpc/h, h)
    "TNG100": (75.0, 0.6774),
    "TNG100_2": (75.0, 0.6774),
    "TNG300": (205.0, 0.6774),
    "EAGLE": (100.0 * 0.6777, 0.6777),
    "HorizonAGN": (100.0, 0.704),
}


def simulation_specs(name: str):
    """(boxsize, h) of a named simulation; KeyError for unknown names."""
    if name in _BOXES:
        return _BOXES[name]
    if "FLAMINGO" in name:
        if "L1" in name:
            return 1000.0 * 0.681, 0.681
        if "L2p8" in name:
            return 2800.0 * 0.681, 0.681
        raise KeyError("Add an L1 or L2p8 suffix to your simname to specify which boxsize is used")
    if "COLIBRE" in name:
        if "L4" in name:
            return 400.0 * 0.681, 0.681
        if "L2" in name:
            return 200.0 * 0.681, 0.681
        raise KeyError("Add an L4 or L2 suffix to your simname to specify which boxsize is used")
    raise KeyError(
        "Simulation name not recognised. Choose from [TNG100, TNG100_2, TNG300, EAGLE, HorizonAGN, "
        "FLAMINGO_L1, FLAMINGO_L2p8, COLIBRE_L400, COLIBRE_L200].")


class SimInfo:
    """boxsize, L_0p5, h and snap_group as SimInfo.__init__ sets them.

    A simulation name wins over an explicit boxsize (Sim_info.py:70-80);
    snap_group is "Snapshot_{n}/" or "" (Sim_info.py:64-69).
    """

    def __init__(self, sim_name, snapshot, boxsize=None, h=None):
        self.simname = sim_name
        if snapshot is None:
            self.snapshot = None
            self.snap_group = ""
        else:
            self.snapshot = str(snapshot)
            self.snap_group = f"Snapshot_{self.snapshot}/"
        if isinstance(sim_name, str):
            self.boxsize, self.h = simulation_specs(sim_name)
            self.L_0p5 = self.boxsize / 2.0
        else:
            self.boxsize = boxsize
            self.h = h
            self.L_0p5 = None if boxsize is None else boxsize / 2.0
