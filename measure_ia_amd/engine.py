"""Device driver of the pair-counting engine (torch plumbing + the C ABI).

torch owns the device memory of inputs and outputs and supplies the stream;
all counting happens in libmeasure_ia_amd.so.  There is no CPU fallback: if
the library or a GPU is missing, constructing an Engine raises.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _capi, hostprep
from .binning import Thresholds

METHODS = {"plain": _capi.MIA_PLAIN, "fast": _capi.MIA_FAST}


@dataclass
class SampleArrays:
    """One galaxy sample as the box counters prepare it (host or device arrays)."""
    positions: object            # (N_p, 3) float64
    shapes: object               # (N_s, 3) float64
    axis: object                 # (N_s, 2) float64, normalised
    e: object                    # (N_s,)   float64
    weight: Optional[object] = None        # (N_p,) or None (unit)
    weight_shape: Optional[object] = None  # (N_s,) or None (unit)
    boxsize: float = 0.0
    los: int = 2
    jk_per_side: int = 0         # L_subboxes (0: no jackknife)
    jk_sub: float = 0.0          # L_sub = 2*L_0p5/L_subboxes
    periodic: bool = True


@dataclass
class LightconeArrays:
    """One lightcone measurement as the lightcone counters prepare it (host or device arrays).

    chi is the comoving distance (already times h with over_h); cos_dec is
    cos(dec/180*pi) of every position; c2, s2 are -e cos 2phi_axis and
    -e sin 2phi_axis of every shape (None for pair counts only)."""
    ra: object
    dec: object
    chi: object
    cos_dec: object
    ra_s: object
    dec_s: object
    chi_s: object
    weight: Optional[object] = None
    weight_shape: Optional[object] = None
    c2: Optional[object] = None
    s2: Optional[object] = None
    patch: Optional[object] = None         # int patch index in [0, n_patches)
    patch_s: Optional[object] = None
    n_patches: int = 0
    proj_scale: float = 1.0
    rp2_cut: float = 0.0


@dataclass
class CountResult:
    npairs: np.ndarray           # (1+K, n_r, n_2) int64 -- [0] total, [1+i] tally of region i
    dd: np.ndarray               # (1+K, n_r, n_2) float64
    splus: np.ndarray            # (1+K, n_r, n_2) float64: sum w_p w_s e+ (no 1/2R)
    scross: np.ndarray           # (n_r, n_2) float64
    variance: Optional[np.ndarray]
    region_stats: Optional[np.ndarray]   # (K, 4): n_pos, n_shape, sum w_s, sum w_s(1-e^2/2)
    stats: dict = field(default_factory=dict)


def _unit(w) -> bool:
    """Whether a weight array can be dropped (unit weights).

    Host arrays are checked on the host.  A device tensor is taken as given:
    checking it would cost a reduction and a sync per count, and a weighted
    count of unit weights gives the same DD (each product is exactly 1)."""
    if w is None or hostprep.is_unit(w):
        return True
    try:
        import torch
        if isinstance(w, torch.Tensor):
            return False
    except ImportError:
        pass
    return bool(np.all(np.asarray(w) == 1.0))


class Engine:
    """A context of the HIP library on one device."""

    def __init__(self, device: int = 0):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("measure_ia_amd: no HIP device visible (the engine has no CPU path)")
        self.torch = torch
        self.device = device
        self.dev = torch.device("cuda", device)
        self.lib = _capi.load()
        ctx = ctypes.c_void_p()
        rc = self.lib.mia_create(device, ctypes.byref(ctx))
        if rc != 0:
            raise _capi.MiaError(rc, (self.lib.mia_last_error(None) or b"").decode())
        self.ctx = ctx

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.mia_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- helpers -------------------------------------------------------------
    def _dev(self, a, shape=None):
        t = self.torch
        if a is None:
            return None
        if isinstance(a, t.Tensor):
            x = a.to(device=self.dev, dtype=t.float64).contiguous()
        else:
            x = t.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64))).to(self.dev)
        if shape is not None and tuple(x.shape) != shape:
            raise ValueError(f"array shape {tuple(x.shape)} != {shape}")
        return x

    def to_device(self, arrays: dict) -> dict:
        """Device float64 copies of a sample's host arrays (for repeated counts);
        weight arrays of unit weights become None (the unweighted path)."""
        out = {}
        for k, v in arrays.items():
            if k in ("weight", "weight_shape") and _unit(v):
                out[k] = None
            else:
                out[k] = self._dev(v)
        return out

    @staticmethod
    def _bins(th: Thresholds, variance: bool):
        r2 = np.ascontiguousarray(th.r2_edges, dtype=np.float64)
        x2 = np.ascontiguousarray(th.x2_edges, dtype=np.float64)
        b = _capi.MiaBins(mode=_capi.MIA_RPPI if th.mode == "rppi" else _capi.MIA_RMU,
                          n_r=th.n_r, n_2=th.n_2, accumulate_variance=1 if variance else 0,
                          r2_edges=r2.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                          x2_edges=x2.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return b, (r2, x2)

    @staticmethod
    def _tuning(tuning):
        if not tuning:
            return None
        t = _capi.MiaTuning()
        for k, v in tuning.items():
            setattr(t, k, int(v))
        return t

    def _prepare(self, s: SampleArrays):
        n_p = int(s.positions.shape[0])
        n_s = int(s.shapes.shape[0])
        pos = self._dev(s.positions, (n_p, 3))
        shp = self._dev(s.shapes, (n_s, 3))
        axis = self._dev(s.axis, (n_s, 2))
        e = self._dev(s.e, (n_s,))
        wp = None if _unit(s.weight) else self._dev(s.weight, (n_p,))
        ws = None if _unit(s.weight_shape) else self._dev(s.weight_shape, (n_s,))
        keep = [pos, shp, axis, e, wp, ws]

        def ptr(x):
            return None if x is None else x.data_ptr()

        ms = _capi.MiaSample(pos=ptr(pos), pos_w=ptr(wp), shp=ptr(shp), shp_w=ptr(ws),
                             axis=ptr(axis), e=ptr(e), n_pos=n_p, n_shape=n_s,
                             boxsize=float(s.boxsize), jk_sub=float(s.jk_sub), los=int(s.los),
                             periodic=1 if s.periodic else 0, jk_per_side=int(s.jk_per_side))
        return ms, keep, (wp is not None or ws is not None)

    def _outputs(self, K, nb, variance):
        """All outputs in one uninitialised device buffer (the count call zeroes
        and fills every one of them); returns (layout, buffer, MiaCounts)."""
        t = self.torch
        sizes = [("npairs", (1 + K) * nb), ("dd", (1 + K) * nb), ("splus", (1 + K) * nb),
                 ("scross", nb), ("variance", nb if variance else 0),
                 ("region_stats", 4 * K)]
        total = sum(n for _, n in sizes)
        buf = t.empty(max(total, 1), dtype=t.float64, device=self.dev)
        layout, off, ptrs = {}, 0, {}
        for name, n in sizes:
            layout[name] = (off, n) if n else None
            ptrs[name] = buf.data_ptr() + 8 * off if n else None
            off += n
        return layout, buf, _capi.MiaCounts(**ptrs)

    @staticmethod
    def _host(o, K, th: Thresholds, stats):
        layout, buf = o
        host = buf.cpu().numpy()  # one device-to-host copy
        shp3 = (1 + K, th.n_r, th.n_2)

        def part(name, shape, dtype=None):
            if layout[name] is None:
                return None
            off, n = layout[name]
            a = host[off:off + n]
            return (a.view(dtype) if dtype else a.copy()).reshape(shape).copy()

        return CountResult(
            npairs=part("npairs", shp3, np.int64),
            dd=part("dd", shp3),
            splus=part("splus", shp3),
            scross=part("scross", (th.n_r, th.n_2)),
            variance=part("variance", (th.n_r, th.n_2)),
            region_stats=part("region_stats", (K, 4)),
            stats=stats)

    # -- API -----------------------------------------------------------------
    def count(self, sample: SampleArrays, th: Thresholds, method: str = "fast",
              variance: bool = False, tuning: Optional[dict] = None) -> CountResult:
        """Count one sample; returns host arrays."""
        ms, keep, _ = self._prepare(sample)
        K = sample.jk_per_side ** 3 if sample.jk_per_side > 0 else 0
        nb = th.n_r * th.n_2
        bins, keep_b = self._bins(th, variance)
        layout, buf, c = self._outputs(K, nb, variance)
        st = _capi.MiaStats()
        tu = self._tuning(tuning)
        stream = self.torch.cuda.current_stream(self.dev).cuda_stream
        rc = self.lib.mia_count(self.ctx, ctypes.byref(ms), ctypes.byref(bins), ctypes.byref(c),
                                METHODS[method], ctypes.byref(tu) if tu else None,
                                ctypes.byref(st), ctypes.c_void_p(stream))
        _capi.check(self.lib, self.ctx, rc)
        return self._host((layout, buf), K, th, st.as_dict())

    def count_batch(self, samples: Sequence[SampleArrays], th: Thresholds,
                    variance: bool = False, tuning: Optional[dict] = None):
        """Count many samples in one persistent launch; returns a list of results."""
        n = len(samples)
        mss = (_capi.MiaSample * n)()
        cs = (_capi.MiaCounts * n)()
        keep, outs, Ks = [], [], []
        nb = th.n_r * th.n_2
        for i, s in enumerate(samples):
            ms, k, _ = self._prepare(s)
            keep.append(k)
            mss[i] = ms
            K = s.jk_per_side ** 3 if s.jk_per_side > 0 else 0
            layout, buf, c = self._outputs(K, nb, variance)
            outs.append((layout, buf))
            Ks.append(K)
            cs[i] = c
        bins, keep_b = self._bins(th, variance)
        st = _capi.MiaStats()
        tu = self._tuning(tuning)
        stream = self.torch.cuda.current_stream(self.dev).cuda_stream
        rc = self.lib.mia_count_batch(self.ctx, n, mss, ctypes.byref(bins), cs,
                                      ctypes.byref(tu) if tu else None, ctypes.byref(st),
                                      ctypes.c_void_p(stream))
        _capi.check(self.lib, self.ctx, rc)
        stats = st.as_dict()
        return [self._host(o, K, th, stats) for o, K in zip(outs, Ks)]

    def prepare(self, sample: SampleArrays, th: Thresholds, tuning: Optional[dict] = None) -> "Prepared":
        """Sort one sample once on the device (mia_prepare_sample); counts of the
        returned handle skip the sort.  The cell grid is chosen for th."""
        ms, keep, weighted = self._prepare(sample)
        bins, keep_b = self._bins(th, False)
        tu = self._tuning(tuning)
        h = ctypes.c_void_p()
        stream = self.torch.cuda.current_stream(self.dev).cuda_stream
        rc = self.lib.mia_prepare_sample(self.ctx, ctypes.byref(ms), ctypes.byref(bins),
                                         ctypes.byref(tu) if tu else None, ctypes.byref(h),
                                         ctypes.c_void_p(stream))
        _capi.check(self.lib, self.ctx, rc)
        K = sample.jk_per_side ** 3 if sample.jk_per_side > 0 else 0
        return Prepared(self, h, K)

    def count_prepared(self, prepared: Sequence["Prepared"], th: Thresholds, variance: bool = False,
                       tuning: Optional[dict] = None):
        """Count prepared samples in one launch (mia_count_prepared); returns a list of results."""
        n = len(prepared)
        hs = (ctypes.c_void_p * n)(*[p.handle for p in prepared])
        cs = (_capi.MiaCounts * n)()
        outs = []
        nb = th.n_r * th.n_2
        for i, p in enumerate(prepared):
            layout, buf, c = self._outputs(p.K, nb, variance)
            outs.append((layout, buf))
            cs[i] = c
        bins, keep_b = self._bins(th, variance)
        st = _capi.MiaStats()
        tu = self._tuning(tuning)
        stream = self.torch.cuda.current_stream(self.dev).cuda_stream
        rc = self.lib.mia_count_prepared(self.ctx, n, hs, ctypes.byref(bins), cs,
                                         ctypes.byref(tu) if tu else None, ctypes.byref(st),
                                         ctypes.c_void_p(stream))
        _capi.check(self.lib, self.ctx, rc)
        stats = st.as_dict()
        return [self._host(o, p.K, th, stats) for o, p in zip(outs, prepared)]

    def _dev_i32(self, a):
        t = self.torch
        if a is None:
            return None
        if isinstance(a, t.Tensor):
            return a.to(device=self.dev, dtype=t.int32).contiguous()
        return t.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.int32))).to(self.dev)

    def count_lightcone(self, lc: LightconeArrays, th: Thresholds,
                        tuning: Optional[dict] = None) -> CountResult:
        """Count one lightcone measurement; returns host arrays with K = n_patches
        tallies (scross is tallied like splus)."""
        n_p = int(np.shape(lc.ra)[0])
        n_s = int(np.shape(lc.ra_s)[0])
        shapes = lc.c2 is not None
        arrs = {k: self._dev(getattr(lc, k), (n_p,)) for k in ("ra", "dec", "chi", "cos_dec")}
        arrs.update({k: self._dev(getattr(lc, k), (n_s,)) for k in ("ra_s", "dec_s", "chi_s")})
        arrs["w"] = None if _unit(lc.weight) else self._dev(lc.weight, (n_p,))
        arrs["w_s"] = None if _unit(lc.weight_shape) else self._dev(lc.weight_shape, (n_s,))
        arrs["c2"] = self._dev(lc.c2, (n_s,)) if shapes else None
        arrs["s2"] = self._dev(lc.s2, (n_s,)) if shapes else None
        K = int(lc.n_patches)
        arrs["patch"] = self._dev_i32(lc.patch) if K else None
        arrs["patch_s"] = self._dev_i32(lc.patch_s) if K else None

        def ptr(x):
            return None if x is None else x.data_ptr()

        ms = _capi.MiaLcSample(n_pos=n_p, n_shape=n_s, proj_scale=float(lc.proj_scale),
                               rp2_cut=float(lc.rp2_cut), n_patches=K,
                               **{k: ptr(v) for k, v in arrs.items()})
        nb = th.n_r * th.n_2
        bins, keep_b = self._bins(th, False)
        tt = self.torch
        sizes = [("npairs", (1 + K) * nb), ("dd", (1 + K) * nb),
                 ("splus", (1 + K) * nb if shapes else 0), ("scross", (1 + K) * nb if shapes else 0)]
        buf = tt.empty(max(sum(n for _, n in sizes), 1), dtype=tt.float64, device=self.dev)
        layout, off, ptrs = {}, 0, {}
        for name, n in sizes:
            layout[name] = (off, n) if n else None
            ptrs[name] = buf.data_ptr() + 8 * off if n else None
            off += n
        c = _capi.MiaCounts(**ptrs)
        st = _capi.MiaStats()
        tu = self._tuning(tuning)
        stream = self.torch.cuda.current_stream(self.dev).cuda_stream
        rc = self.lib.mia_count_lightcone(self.ctx, ctypes.byref(ms), ctypes.byref(bins),
                                          ctypes.byref(c), ctypes.byref(tu) if tu else None,
                                          ctypes.byref(st), ctypes.c_void_p(stream))
        _capi.check(self.lib, self.ctx, rc)
        host = buf.cpu().numpy()
        shp3 = (1 + K, th.n_r, th.n_2)

        def part(name, dtype=None):
            if layout[name] is None:
                return None
            o, n = layout[name]
            a = host[o:o + n]
            return (a.view(dtype) if dtype else a).reshape(shp3).copy()

        return CountResult(npairs=part("npairs", np.int64), dd=part("dd"), splus=part("splus"),
                           scross=part("scross"), variance=None, region_stats=None,
                           stats=st.as_dict())

    def region_ids(self, xyz, jk_per_side: int, jk_sub: float) -> np.ndarray:
        x = self._dev(xyz)
        n = int(x.shape[0])
        out = self.torch.zeros(max(n, 1), dtype=self.torch.int32, device=self.dev)
        stream = self.torch.cuda.current_stream(self.dev).cuda_stream
        rc = self.lib.mia_region_ids(self.ctx, x.data_ptr(), n, float(jk_sub), int(jk_per_side),
                                     out.data_ptr(), ctypes.c_void_p(stream))
        _capi.check(self.lib, self.ctx, rc)
        return out[:n].cpu().numpy()


class Prepared:
    """A sample sorted on the device (mia_prepare_sample), released with the object."""

    def __init__(self, engine: Engine, handle, K: int):
        self.engine = engine
        self.handle = handle
        self.K = K

    def release(self):
        if self.handle is not None and self.engine.ctx:
            self.engine.lib.mia_release_prepared(self.engine.ctx, self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


_ENGINES = {}


def get_engine(device: int = 0) -> Engine:
    e = _ENGINES.get(device)
    if e is None:
        e = _ENGINES[device] = Engine(device)
    return e
