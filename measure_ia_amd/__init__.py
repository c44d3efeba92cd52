"""MI355X-native pair-counting engine for MeasureIA.

``MeasureIABox`` is a drop-in for the reference's box API
(src/measureia/measure_IA.py:10-262) and ``MeasureIALightcone`` for its
lightcone API (measure_IA.py:265-1060): the pair loops, the binning and the
jackknife tallies run in hand-written HIP kernels (libmeasure_ia_amd.so,
C ABI in include/measure_ia_amd.h); input preparation, estimators and the
HDF5 output stay on the host.
"""
from .box import MeasureIABox
from .lightcone import MeasureIALightcone
from .readout import ReadData
from .siminfo import SimInfo

__all__ = ["MeasureIABox", "MeasureIALightcone", "ReadData", "SimInfo"]
