"""Minimal HDF5 store for the MeasureIA output layout.

The reference writes its results with h5py through two helpers,
``write_dataset_hdf5`` (delete-and-recreate a dataset) and
``create_group_hdf5`` (mkdir -p of a group path) in
src/measureia/write_data.py:1-56.  h5py is not always installed next to a
ROCm stack, so this module talks to the HDF5 C library directly through
ctypes and falls back to h5py only when the C library cannot be found.

Only what the output layout needs is implemented: float64/int64 datasets of
any rank, nested groups, listing, deleting, and reading back.  One
``H5Store`` keeps the file open for a whole measurement, instead of the
reference's open/close per post-processing call.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import os
from typing import Dict, Iterator, List, Optional

import numpy as np

__all__ = ["H5Store", "hdf5_backend", "read_dataset", "list_tree"]

_hid = ctypes.c_int64
_herr = ctypes.c_int
_hsize = ctypes.c_ulonglong

_H5F_ACC_RDONLY = 0x0000
_H5F_ACC_RDWR = 0x0001
_H5F_ACC_TRUNC = 0x0002
_H5P_DEFAULT = 0
_H5S_ALL = 0
_H5I_GROUP = 2
_H5I_DATASET = 5
_H5_INDEX_NAME = 0
_H5_ITER_INC = 0
_H5T_FLOAT = 1
_H5T_INTEGER = 0

_ITER_CB = ctypes.CFUNCTYPE(_herr, _hid, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p)

_lib = None
_lib_error: Optional[str] = None


def _candidate_paths() -> List[str]:
    env = os.environ.get("MEASURE_IA_HDF5_LIB")
    paths = [env] if env else []
    found = ctypes.util.find_library("hdf5")
    if found:
        paths.append(found)
    paths += ["/opt/conda/lib/libhdf5.so", "/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5.so",
              "/usr/lib/x86_64-linux-gnu/libhdf5_serial.so", "/usr/lib64/libhdf5.so"]
    return paths


def _load():
    global _lib, _lib_error
    if _lib is not None or _lib_error is not None:
        return _lib
    for p in _candidate_paths():
        try:
            lib = ctypes.CDLL(p)
        except OSError:
            continue
        _bind(lib)
        if lib.H5open() < 0:
            continue
        # Errors are reported through return codes and raised as Python
        # exceptions; HDF5's own stderr trace is switched off.
        lib.H5Eset_auto2(0, None, None)
        _lib = lib
        return _lib
    _lib_error = "libhdf5 not found (set MEASURE_IA_HDF5_LIB)"
    return None


def _bind(lib):
    def f(name, res, *args):
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = list(args)
    f("H5open", _herr)
    f("H5Eset_auto2", _herr, _hid, ctypes.c_void_p, ctypes.c_void_p)
    f("H5Fcreate", _hid, ctypes.c_char_p, ctypes.c_uint, _hid, _hid)
    f("H5Fopen", _hid, ctypes.c_char_p, ctypes.c_uint, _hid)
    f("H5Fclose", _herr, _hid)
    f("H5Fflush", _herr, _hid, ctypes.c_int)
    f("H5Gcreate2", _hid, _hid, ctypes.c_char_p, _hid, _hid, _hid)
    f("H5Gopen2", _hid, _hid, ctypes.c_char_p, _hid)
    f("H5Gclose", _herr, _hid)
    f("H5Dcreate2", _hid, _hid, ctypes.c_char_p, _hid, _hid, _hid, _hid, _hid)
    f("H5Dopen2", _hid, _hid, ctypes.c_char_p, _hid)
    f("H5Dclose", _herr, _hid)
    f("H5Dwrite", _herr, _hid, _hid, _hid, _hid, _hid, ctypes.c_void_p)
    f("H5Dread", _herr, _hid, _hid, _hid, _hid, _hid, ctypes.c_void_p)
    f("H5Dget_space", _hid, _hid)
    f("H5Dget_type", _hid, _hid)
    f("H5Tget_class", ctypes.c_int, _hid)
    f("H5Tclose", _herr, _hid)
    f("H5Screate_simple", _hid, ctypes.c_int, ctypes.POINTER(_hsize), ctypes.POINTER(_hsize))
    f("H5Screate", _hid, ctypes.c_int)
    f("H5Sget_simple_extent_ndims", ctypes.c_int, _hid)
    f("H5Sget_simple_extent_dims", ctypes.c_int, _hid, ctypes.POINTER(_hsize), ctypes.POINTER(_hsize))
    f("H5Sclose", _herr, _hid)
    f("H5Lexists", ctypes.c_int, _hid, ctypes.c_char_p, _hid)
    f("H5Ldelete", _herr, _hid, ctypes.c_char_p, _hid)
    f("H5Literate", _herr, _hid, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_hsize), _ITER_CB,
      ctypes.c_void_p)
    f("H5Oopen", _hid, _hid, ctypes.c_char_p, _hid)
    f("H5Oclose", _herr, _hid)
    f("H5Iget_type", ctypes.c_int, _hid)
    for g in ("H5T_NATIVE_DOUBLE_g", "H5T_NATIVE_INT64_g"):
        setattr(lib, "_" + g, _hid.in_dll(lib, g))


def hdf5_backend() -> str:
    """Name of the HDF5 backend in use: 'libhdf5', 'h5py' or 'none'."""
    if _load() is not None:
        return "libhdf5"
    try:
        import h5py  # noqa: F401
        return "h5py"
    except ImportError:
        return "none"


def _norm(path: str) -> str:
    """Collapse '//' and strip slashes: HDF5 treats 'a//b' as 'a/b'."""
    parts = [p for p in path.split("/") if p]
    return "/".join(parts)


class H5Store:
    """An open HDF5 file with the two operations the output layout needs.

    mode 'a' opens read/write and creates a missing file (h5py's 'a');
    'r' opens read-only; 'w' truncates.
    """

    def __init__(self, filename: str, mode: str = "a"):
        self.filename = str(filename)
        self.mode = mode
        self._h5py = None
        # groups known to exist, and the datasets this store wrote (served
        # back by read() without a file round trip; the file holds the same)
        self._groups = {""}
        self._written = {}
        # groups this store created (a dataset in one exists only if this
        # store wrote it) and the dataspaces of the shapes it wrote
        self._fresh = set()
        self._spaces = {}
        lib = _load()
        if lib is None:
            try:
                import h5py
            except ImportError as exc:  # pragma: no cover - depends on host
                raise RuntimeError(f"no HDF5 backend: {_lib_error}; h5py missing") from exc
            self._h5py = h5py.File(self.filename, mode)
            return
        self._lib = lib
        name = self.filename.encode()
        if mode == "r":
            fid = lib.H5Fopen(name, _H5F_ACC_RDONLY, _H5P_DEFAULT)
        elif mode == "w":
            fid = lib.H5Fcreate(name, _H5F_ACC_TRUNC, _H5P_DEFAULT, _H5P_DEFAULT)
        elif mode == "a":
            if os.path.exists(self.filename):
                fid = lib.H5Fopen(name, _H5F_ACC_RDWR, _H5P_DEFAULT)
            else:
                fid = lib.H5Fcreate(name, _H5F_ACC_TRUNC, _H5P_DEFAULT, _H5P_DEFAULT)
        else:
            raise ValueError(f"unsupported mode {mode!r}")
        if fid < 0:
            raise OSError(f"cannot open HDF5 file {self.filename!r} (mode {mode!r})")
        self._fid = fid

    # -- context management -------------------------------------------------
    def close(self):
        if self._h5py is not None:
            self._h5py.close()
            self._h5py = None
            return
        if getattr(self, "_fid", -1) >= 0:
            for sid in self._spaces.values():
                self._lib.H5Sclose(sid)
            self._spaces = {}
            self._lib.H5Fclose(self._fid)
            self._fid = -1

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- queries -------------------------------------------------------------
    def exists(self, path: str) -> bool:
        path = _norm(path)
        if path == "":
            return True
        if self._h5py is not None:
            return path in self._h5py
        # H5Lexists requires every intermediate link to exist.
        parts = path.split("/")
        for i in range(1, len(parts) + 1):
            sub = "/".join(parts[:i]).encode()
            if self._lib.H5Lexists(self._fid, sub, _H5P_DEFAULT) <= 0:
                return False
        return True

    def kind(self, path: str) -> Optional[str]:
        """'group', 'dataset' or None."""
        path = _norm(path)
        if not self.exists(path):
            return None
        if path == "":
            return "group"
        if self._h5py is not None:
            import h5py
            return "group" if isinstance(self._h5py[path], h5py.Group) else "dataset"
        oid = self._lib.H5Oopen(self._fid, path.encode(), _H5P_DEFAULT)
        if oid < 0:
            return None
        t = self._lib.H5Iget_type(oid)
        self._lib.H5Oclose(oid)
        return {_H5I_GROUP: "group", _H5I_DATASET: "dataset"}.get(t)

    def keys(self, group: str = "") -> List[str]:
        group = _norm(group)
        if self._h5py is not None:
            g = self._h5py[group] if group else self._h5py
            return list(g.keys())
        gid = self._lib.H5Gopen2(self._fid, (group or "/").encode(), _H5P_DEFAULT)
        if gid < 0:
            raise KeyError(group)
        names: List[str] = []

        def cb(_g, name, _info, _data):
            names.append(name.decode())
            return 0

        cfun = _ITER_CB(cb)
        idx = _hsize(0)
        rc = self._lib.H5Literate(gid, _H5_INDEX_NAME, _H5_ITER_INC, ctypes.byref(idx), cfun, None)
        self._lib.H5Gclose(gid)
        if rc < 0:
            raise OSError(f"cannot list group {group!r}")
        return names

    def walk(self, group: str = "") -> Iterator[str]:
        """Yield every dataset path below ``group``."""
        for k in self.keys(group):
            p = f"{_norm(group)}/{k}" if _norm(group) else k
            if self.kind(p) == "group":
                yield from self.walk(p)
            else:
                yield p

    # -- groups --------------------------------------------------------------
    def require_group(self, path: str) -> str:
        """mkdir -p of a group path (write_data.py:23-56); returns the path."""
        path = _norm(path)
        if path in self._groups:
            return path
        if self._h5py is not None:
            self._h5py.require_group(path)
            self._groups.add(path)
            return path
        parts = path.split("/")
        for i in range(1, len(parts) + 1):
            sub = "/".join(parts[:i])
            if self._lib.H5Lexists(self._fid, sub.encode(), _H5P_DEFAULT) > 0:
                continue
            gid = self._lib.H5Gcreate2(self._fid, sub.encode(), _H5P_DEFAULT, _H5P_DEFAULT,
                                       _H5P_DEFAULT)
            if gid < 0:
                raise OSError(f"cannot create group {sub!r}")
            self._lib.H5Gclose(gid)
            self._fresh.add(sub)
        for i in range(1, len(parts) + 1):
            self._groups.add("/".join(parts[:i]))
        return path

    # -- datasets ------------------------------------------------------------
    def write(self, group: str, name: str, data) -> None:
        """Overwrite-or-create a dataset (write_data.py:1-20)."""
        group = self.require_group(group)
        path = f"{group}/{name}" if group else name
        path = _norm(path)
        arr = np.asarray(data)
        if arr.dtype.kind in "iub":
            arr = np.ascontiguousarray(arr, dtype=np.int64)
        else:
            arr = np.ascontiguousarray(arr, dtype=np.float64)
        seen = path in self._written
        self._written[path] = arr.copy()
        if self._h5py is not None:
            if path in self._h5py:
                del self._h5py[path]
            self._h5py.create_dataset(path, data=arr)
            return
        lib = self._lib
        bpath = path.encode()
        if (seen or group not in self._fresh) and lib.H5Lexists(self._fid, bpath, _H5P_DEFAULT) > 0:
            if lib.H5Ldelete(self._fid, bpath, _H5P_DEFAULT) < 0:
                raise OSError(f"cannot delete {path!r}")
        sid = self._spaces.get(arr.shape)
        if sid is None:
            if arr.ndim == 0:
                sid = lib.H5Screate(0)  # H5S_SCALAR
            else:
                dims = (_hsize * arr.ndim)(*arr.shape)
                sid = lib.H5Screate_simple(arr.ndim, dims, None)
            if sid < 0:
                raise OSError("H5Screate failed")
            self._spaces[arr.shape] = sid
        mtype = lib._H5T_NATIVE_INT64_g.value if arr.dtype == np.int64 else lib._H5T_NATIVE_DOUBLE_g.value
        did = lib.H5Dcreate2(self._fid, bpath, mtype, sid, _H5P_DEFAULT, _H5P_DEFAULT, _H5P_DEFAULT)
        if did < 0:
            raise OSError(f"cannot create dataset {path!r}")
        rc = lib.H5Dwrite(did, mtype, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT,
                          arr.ctypes.data_as(ctypes.c_void_p))
        lib.H5Dclose(did)
        if rc < 0:
            raise OSError(f"cannot write dataset {path!r}")

    def read(self, path: str) -> np.ndarray:
        path = _norm(path)
        if path in self._written:
            return self._written[path].copy()
        if self._h5py is not None:
            return np.asarray(self._h5py[path][()])
        lib = self._lib
        did = lib.H5Dopen2(self._fid, path.encode(), _H5P_DEFAULT)
        if did < 0:
            raise KeyError(path)
        try:
            sid = lib.H5Dget_space(did)
            nd = lib.H5Sget_simple_extent_ndims(sid)
            dims = (_hsize * max(nd, 1))()
            if nd > 0:
                lib.H5Sget_simple_extent_dims(sid, dims, None)
            lib.H5Sclose(sid)
            shape = tuple(int(dims[i]) for i in range(nd))
            tid = lib.H5Dget_type(did)
            cls = lib.H5Tget_class(tid)
            lib.H5Tclose(tid)
            if cls == _H5T_INTEGER:
                out = np.empty(shape, dtype=np.int64)
                mtype = lib._H5T_NATIVE_INT64_g.value
            else:
                out = np.empty(shape, dtype=np.float64)
                mtype = lib._H5T_NATIVE_DOUBLE_g.value
            rc = lib.H5Dread(did, mtype, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT,
                             out.ctypes.data_as(ctypes.c_void_p))
            if rc < 0:
                raise OSError(f"cannot read {path!r}")
            return out
        finally:
            lib.H5Dclose(did)

    def flush(self):
        if self._h5py is not None:
            self._h5py.flush()
        else:
            self._lib.H5Fflush(self._fid, 1)


def read_dataset(filename: str, path: str) -> np.ndarray:
    with H5Store(filename, "r") as f:
        return f.read(path)


def list_tree(filename: str) -> Dict[str, tuple]:
    """Every dataset path in a file mapped to its shape."""
    out = {}
    with H5Store(filename, "r") as f:
        for p in f.walk(""):
            out[p] = f.read(p).shape
    return out
