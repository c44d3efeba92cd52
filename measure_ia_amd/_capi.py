"""ctypes binding of include/measure_ia_amd.h.

The structures below mirror the header field for field.  The library is
loaded after torch, so that it shares torch's HIP runtime (same soname) and
device pointers and streams from torch tensors can be passed straight in.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# MEASURE_IA_LIB overrides the library path (e.g. an alternative build).
LIB_PATH = os.environ.get("MEASURE_IA_LIB") or os.path.join(HERE, "libmeasure_ia_amd.so")

MIA_RPPI, MIA_RMU = 0, 1
MIA_PLAIN, MIA_FAST = 0, 1
STATUS = {0: "MIA_OK", -1: "MIA_E_ARG", -2: "MIA_E_HIP", -3: "MIA_E_NOMEM", -4: "MIA_E_RANGE",
          -5: "MIA_E_UNSUPPORTED"}

_dp = ctypes.POINTER(ctypes.c_double)


class MiaSample(ctypes.Structure):
    _fields_ = [("pos", ctypes.c_void_p), ("pos_w", ctypes.c_void_p),
                ("shp", ctypes.c_void_p), ("shp_w", ctypes.c_void_p),
                ("axis", ctypes.c_void_p), ("e", ctypes.c_void_p),
                ("n_pos", ctypes.c_int64), ("n_shape", ctypes.c_int64),
                ("boxsize", ctypes.c_double), ("jk_sub", ctypes.c_double),
                ("los", ctypes.c_int32), ("periodic", ctypes.c_int32),
                ("jk_per_side", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class MiaLcSample(ctypes.Structure):
    _fields_ = [("ra", ctypes.c_void_p), ("dec", ctypes.c_void_p), ("chi", ctypes.c_void_p),
                ("cos_dec", ctypes.c_void_p), ("w", ctypes.c_void_p), ("patch", ctypes.c_void_p),
                ("ra_s", ctypes.c_void_p), ("dec_s", ctypes.c_void_p), ("chi_s", ctypes.c_void_p),
                ("w_s", ctypes.c_void_p), ("c2", ctypes.c_void_p), ("s2", ctypes.c_void_p),
                ("patch_s", ctypes.c_void_p),
                ("n_pos", ctypes.c_int64), ("n_shape", ctypes.c_int64),
                ("proj_scale", ctypes.c_double), ("rp2_cut", ctypes.c_double),
                ("n_patches", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class MiaBins(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("n_r", ctypes.c_int32), ("n_2", ctypes.c_int32),
                ("accumulate_variance", ctypes.c_int32),
                ("r2_edges", _dp), ("x2_edges", _dp)]


class MiaCounts(ctypes.Structure):
    _fields_ = [("npairs", ctypes.c_void_p), ("dd", ctypes.c_void_p), ("splus", ctypes.c_void_p),
                ("scross", ctypes.c_void_p), ("variance", ctypes.c_void_p),
                ("region_stats", ctypes.c_void_p)]


class MiaTuning(ctypes.Structure):
    _fields_ = [("columns_per_side", ctypes.c_int32), ("slabs_per_side", ctypes.c_int32),
                ("slabs_per_block", ctypes.c_int32), ("block_shapes", ctypes.c_int32),
                ("item_positions", ctypes.c_int32), ("waves", ctypes.c_int32),
                ("hist_copies", ctypes.c_int32), ("cell_shapes", ctypes.c_int32),
                ("r_bins_per_pass", ctypes.c_int32), ("graph", ctypes.c_int32),
                ("deterministic", ctypes.c_int32)]


class MiaStats(ctypes.Structure):
    _fields_ = [("ms_prepare", ctypes.c_double), ("ms_plan", ctypes.c_double),
                ("ms_pairs", ctypes.c_double), ("ms_finalize", ctypes.c_double),
                ("candidates", ctypes.c_int64), ("items", ctypes.c_int64),
                ("rechecks", ctypes.c_int64),
                ("columns_per_side", ctypes.c_int32), ("slabs_per_side", ctypes.c_int32),
                ("slabs_per_block", ctypes.c_int32), ("block_shapes", ctypes.c_int32),
                ("waves", ctypes.c_int32), ("method", ctypes.c_int32),
                ("ms_total", ctypes.c_double), ("graph", ctypes.c_int32), ("deterministic", ctypes.c_int32),
                ("item_capacity", ctypes.c_int64), ("wave_iterations", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# name -> (restype, argtypes): the complete exported surface of the header.
SIGNATURES = {
    "mia_abi_version": (ctypes.c_int, []),
    "mia_create": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]),
    "mia_destroy": (None, [ctypes.c_void_p]),
    "mia_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "mia_count": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(MiaSample), ctypes.POINTER(MiaBins),
                                 ctypes.POINTER(MiaCounts), ctypes.c_int32,
                                 ctypes.POINTER(MiaTuning), ctypes.POINTER(MiaStats),
                                 ctypes.c_void_p]),
    "mia_count_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(MiaSample),
                                       ctypes.POINTER(MiaBins), ctypes.POINTER(MiaCounts),
                                       ctypes.POINTER(MiaTuning), ctypes.POINTER(MiaStats),
                                       ctypes.c_void_p]),
    "mia_count_lightcone": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(MiaLcSample),
                                           ctypes.POINTER(MiaBins), ctypes.POINTER(MiaCounts),
                                           ctypes.POINTER(MiaTuning), ctypes.POINTER(MiaStats),
                                           ctypes.c_void_p]),
    "mia_region_ids": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                      ctypes.c_double, ctypes.c_int32, ctypes.c_void_p,
                                      ctypes.c_void_p]),
    "mia_count_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(MiaSample), ctypes.POINTER(MiaBins),
                                         ctypes.POINTER(MiaCounts), ctypes.POINTER(MiaTuning), ctypes.c_void_p]),
    "mia_noalloc_workspace_bytes": (ctypes.c_int64, []),
    "mia_prepare_sample": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(MiaSample), ctypes.POINTER(MiaBins),
                                          ctypes.POINTER(MiaTuning), ctypes.POINTER(ctypes.c_void_p),
                                          ctypes.c_void_p]),
    "mia_count_prepared": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p),
                                          ctypes.POINTER(MiaBins), ctypes.POINTER(MiaCounts),
                                          ctypes.POINTER(MiaTuning), ctypes.POINTER(MiaStats),
                                          ctypes.c_void_p]),
    "mia_release_prepared": (None, [ctypes.c_void_p, ctypes.c_void_p]),
}

_lib = None


class LibraryMissing(RuntimeError):
    pass


def load(path: str = LIB_PATH):
    """Load the HIP library (raises LibraryMissing if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise LibraryMissing(f"{path} not found: build it with `python -m measure_ia_amd.build`")
    try:
        import torch  # noqa: F401  (share torch's HIP runtime)
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class MiaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


def check(lib, ctx, rc):
    if rc != 0:
        msg = lib.mia_last_error(ctx)
        raise MiaError(rc, msg.decode() if msg else "")
