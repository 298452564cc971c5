"""CXI (HDF5) reading and writing for patterns in and reconstructions out.

Provides the interface of the reference's ``spimage.CXIWriter`` and
``spimage.CXIReader`` (src/_spimage_cxi.py:5-292): the same constructors,
methods, keyword arguments and event dictionaries.  Configuration errors raise
ValueError instead of ending the process.  The reference builds on h5py, which this image does not
have; this module binds the HDF5 C library (libhdf5, 1.10 API) directly
with ctypes.

Types: numpy integers and floats map to the native HDF5 types; complex
arrays are stored as the compound {r, i} that h5py writes; booleans are
stored as uint8 (h5py writes an enum).  Scalar datasets and the "axes"
attribute are written as the reference writes them.  String datasets are not
supported.
"""
from __future__ import annotations

import ctypes as C
import ctypes.util
import os
import sys

import numpy as np

__all__ = ["CXIWriter", "CXIReader", "H5File"]

_hid = C.c_int64
_herr = C.c_int
_hsize = C.c_ulonglong
H5P_DEFAULT = 0
H5F_ACC_RDONLY, H5F_ACC_RDWR, H5F_ACC_TRUNC = 0, 1, 2
H5S_ALL, H5S_SCALAR, H5S_SELECT_SET = 0, 0, 0
H5T_INTEGER, H5T_FLOAT, H5T_STRING, H5T_COMPOUND, H5T_ENUM = 0, 1, 3, 6, 8
H5T_SGN_NONE = 0


def _find_lib():
    cands = [os.environ.get("SPX_HDF5_LIB", ""), ctypes.util.find_library("hdf5") or "",
             os.path.join(sys.prefix, "lib", "libhdf5.so"), "/opt/conda/lib/libhdf5.so", "libhdf5.so"]
    for c in cands:
        if not c:
            continue
        try:
            return C.CDLL(c)
        except OSError:
            continue
    raise OSError("libhdf5 not found (set SPX_HDF5_LIB to its path)")


class _H5:
    """The libhdf5 entry points this module uses."""

    def __init__(self):
        L = self.lib = _find_lib()
        sig = {
            "H5open": (_herr, []),
            "H5Fcreate": (_hid, [C.c_char_p, C.c_uint, _hid, _hid]),
            "H5Fopen": (_hid, [C.c_char_p, C.c_uint, _hid]),
            "H5Fclose": (_herr, [_hid]),
            "H5Gcreate2": (_hid, [_hid, C.c_char_p, _hid, _hid, _hid]),
            "H5Gclose": (_herr, [_hid]),
            "H5Lexists": (C.c_int, [_hid, C.c_char_p, _hid]),
            "H5Dcreate2": (_hid, [_hid, C.c_char_p, _hid, _hid, _hid, _hid, _hid]),
            "H5Dopen2": (_hid, [_hid, C.c_char_p, _hid]),
            "H5Dclose": (_herr, [_hid]),
            "H5Dget_space": (_hid, [_hid]),
            "H5Dget_type": (_hid, [_hid]),
            "H5Dread": (_herr, [_hid, _hid, _hid, _hid, _hid, C.c_void_p]),
            "H5Dwrite": (_herr, [_hid, _hid, _hid, _hid, _hid, C.c_void_p]),
            "H5Screate": (_hid, [C.c_int]),
            "H5Screate_simple": (_hid, [C.c_int, C.POINTER(_hsize), C.POINTER(_hsize)]),
            "H5Sclose": (_herr, [_hid]),
            "H5Sget_simple_extent_ndims": (C.c_int, [_hid]),
            "H5Sget_simple_extent_dims": (C.c_int, [_hid, C.POINTER(_hsize), C.POINTER(_hsize)]),
            "H5Sselect_hyperslab": (_herr, [_hid, C.c_int, C.POINTER(_hsize), C.POINTER(_hsize),
                                            C.POINTER(_hsize), C.POINTER(_hsize)]),
            "H5Tcopy": (_hid, [_hid]),
            "H5Tclose": (_herr, [_hid]),
            "H5Tcreate": (_hid, [C.c_int, C.c_size_t]),
            "H5Tinsert": (_herr, [_hid, C.c_char_p, C.c_size_t, _hid]),
            "H5Tset_size": (_herr, [_hid, C.c_size_t]),
            "H5Tget_class": (C.c_int, [_hid]),
            "H5Tget_size": (C.c_size_t, [_hid]),
            "H5Tget_sign": (C.c_int, [_hid]),
            "H5Tget_nmembers": (C.c_int, [_hid]),
            "H5Tget_member_type": (_hid, [_hid, C.c_uint]),
            "H5Tget_member_class": (C.c_int, [_hid, C.c_uint]),
            "H5Acreate2": (_hid, [_hid, C.c_char_p, _hid, _hid, _hid, _hid]),
            "H5Awrite": (_herr, [_hid, _hid, C.c_void_p]),
            "H5Aclose": (_herr, [_hid]),
        }
        for name, s in sig.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = s
        if L.H5open() < 0:
            raise OSError("H5open failed")
        g = lambda n: _hid.in_dll(L, n).value  # noqa: E731
        self.native = {
            np.dtype(np.int8): g("H5T_NATIVE_INT8_g"), np.dtype(np.uint8): g("H5T_NATIVE_UINT8_g"),
            np.dtype(np.int16): g("H5T_NATIVE_INT16_g"), np.dtype(np.uint16): g("H5T_NATIVE_UINT16_g"),
            np.dtype(np.int32): g("H5T_NATIVE_INT32_g"), np.dtype(np.uint32): g("H5T_NATIVE_UINT32_g"),
            np.dtype(np.int64): g("H5T_NATIVE_INT64_g"), np.dtype(np.uint64): g("H5T_NATIVE_UINT64_g"),
            np.dtype(np.float32): g("H5T_NATIVE_FLOAT_g"), np.dtype(np.float64): g("H5T_NATIVE_DOUBLE_g"),
        }
        self.c_s1 = g("H5T_C_S1_g")

    def check(self, rc, what):
        if rc < 0:
            raise OSError(f"HDF5: {what} failed")
        return rc

    def mem_type(self, dt: np.dtype):
        """(type id, owned) for a numpy dtype; complex -> compound {r, i}."""
        dt = np.dtype(dt)
        if dt == np.bool_:
            dt = np.dtype(np.uint8)
        if dt.kind == "c":
            part = np.dtype(np.float32 if dt.itemsize == 8 else np.float64)
            t = self.check(self.lib.H5Tcreate(H5T_COMPOUND, dt.itemsize), "H5Tcreate")
            self.check(self.lib.H5Tinsert(t, b"r", 0, self.native[part]), "H5Tinsert")
            self.check(self.lib.H5Tinsert(t, b"i", part.itemsize, self.native[part]), "H5Tinsert")
            return t, True
        if dt not in self.native:
            raise TypeError(f"CXI: data type {dt} is not supported")
        return self.native[dt], False

    def dtype_of(self, tid) -> np.dtype:
        """numpy dtype that holds a file type."""
        L = self.lib
        cls = L.H5Tget_class(tid)
        size = L.H5Tget_size(tid)
        if cls == H5T_INTEGER:
            signed = L.H5Tget_sign(tid) != H5T_SGN_NONE
            return np.dtype(f"{'i' if signed else 'u'}{size}")
        if cls == H5T_FLOAT:
            return np.dtype(f"f{size}")
        if cls == H5T_COMPOUND and L.H5Tget_nmembers(tid) == 2 and \
                L.H5Tget_member_class(tid, 0) == H5T_FLOAT and L.H5Tget_member_class(tid, 1) == H5T_FLOAT:
            return np.dtype(f"c{size}")
        raise TypeError(f"CXI: HDF5 type class {cls} (size {size}) is not supported")


_h5 = None


def _lib() -> _H5:
    global _h5
    if _h5 is None:
        _h5 = _H5()
    return _h5


def _dims(shape):
    return (_hsize * max(1, len(shape)))(*shape) if shape else None


class H5File:
    """A minimal HDF5 file: groups, fixed-size datasets, slices."""

    def __init__(self, filename, mode="r"):
        h = self.h = _lib()
        name = os.fsencode(filename)
        if mode == "w":
            self.fid = h.check(h.lib.H5Fcreate(name, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), f"create {filename}")
        elif mode in ("r", "r+"):
            flags = H5F_ACC_RDONLY if mode == "r" else H5F_ACC_RDWR
            self.fid = h.check(h.lib.H5Fopen(name, flags, H5P_DEFAULT), f"open {filename}")
        else:
            raise ValueError(mode)

    def close(self):
        if self.fid is not None:
            self.h.lib.H5Fclose(self.fid)
            self.fid = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __contains__(self, path) -> bool:
        parts = [p for p in path.split("/") if p]
        cur = ""
        for p in parts:
            cur += "/" + p
            if self.h.lib.H5Lexists(self.fid, cur.encode(), H5P_DEFAULT) <= 0:
                return False
        return True

    def create_group(self, path):
        parts = [p for p in path.split("/") if p]
        cur = ""
        for p in parts:
            cur += "/" + p
            if cur not in self:
                g = self.h.check(self.h.lib.H5Gcreate2(self.fid, cur.encode(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                                 f"create group {cur}")
                self.h.lib.H5Gclose(g)

    def create_dataset(self, path, shape, dtype, axes=None):
        h, L = self.h, self.h.lib
        parent = path.rsplit("/", 1)[0]
        if parent:
            self.create_group(parent)
        tid, own = h.mem_type(dtype)
        sid = h.check(L.H5Screate_simple(len(shape), _dims(shape), None), "H5Screate_simple")
        try:
            did = h.check(L.H5Dcreate2(self.fid, path.encode(), tid, sid, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                          f"create dataset {path}")
            if axes is not None:
                self._axes(did, axes)
            L.H5Dclose(did)
        finally:
            L.H5Sclose(sid)
            if own:
                L.H5Tclose(tid)

    def _axes(self, did, axes):
        # one fixed-length string, stored as an array of one (like the list the reference writes)
        h, L = self.h, self.h.lib
        raw = axes.encode()
        st = h.check(L.H5Tcopy(h.c_s1), "H5Tcopy")
        L.H5Tset_size(st, len(raw))
        sid = h.check(L.H5Screate_simple(1, _dims((1,)), None), "H5Screate_simple")
        aid = h.check(L.H5Acreate2(did, b"axes", st, sid, H5P_DEFAULT, H5P_DEFAULT), "H5Acreate2")
        buf = C.create_string_buffer(raw, len(raw))
        L.H5Awrite(aid, st, buf)
        L.H5Aclose(aid)
        L.H5Sclose(sid)
        L.H5Tclose(st)

    def shape(self, path):
        h, L = self.h, self.h.lib
        did = h.check(L.H5Dopen2(self.fid, path.encode(), H5P_DEFAULT), f"open dataset {path}")
        sid = L.H5Dget_space(did)
        nd = L.H5Sget_simple_extent_ndims(sid)
        dims = (_hsize * max(nd, 1))()
        L.H5Sget_simple_extent_dims(sid, dims, None)
        L.H5Sclose(sid)
        L.H5Dclose(did)
        return tuple(int(dims[i]) for i in range(nd))

    def _io(self, path, index, data=None):
        """Read (data None) or write dataset `path`, whole (index None) or
        its index-th entry along the first axis."""
        h, L = self.h, self.h.lib
        did = h.check(L.H5Dopen2(self.fid, path.encode(), H5P_DEFAULT), f"open dataset {path}")
        fsid = L.H5Dget_space(did)
        ftid = L.H5Dget_type(did)
        try:
            nd = L.H5Sget_simple_extent_ndims(fsid)
            dims = (_hsize * max(nd, 1))()
            L.H5Sget_simple_extent_dims(fsid, dims, None)
            full = tuple(int(dims[i]) for i in range(nd))
            dt = h.dtype_of(ftid) if data is None else np.asarray(data).dtype
            if index is None:
                shape, msid = full, H5S_ALL
                sel = H5S_ALL
            else:
                if nd == 0 or not 0 <= index < full[0]:
                    raise IndexError(f"{path}: index {index} outside {full}")
                shape = full[1:]
                start = (_hsize * nd)(*([index] + [0] * (nd - 1)))
                count = (_hsize * nd)(*([1] + list(full[1:])))
                h.check(L.H5Sselect_hyperslab(fsid, H5S_SELECT_SET, start, None, count, None), "select")
                sel = fsid
                msid = h.check(L.H5Screate_simple(len(shape), _dims(shape), None) if shape
                               else L.H5Screate(H5S_SCALAR), "memory space")
            tid, own = h.mem_type(dt)
            try:
                if data is None:
                    out = np.empty(shape, dt)
                    h.check(L.H5Dread(did, tid, msid, sel, H5P_DEFAULT, out.ctypes.data), f"read {path}")
                    return out
                arr = np.ascontiguousarray(data, dtype=np.uint8 if dt == np.bool_ else dt).reshape(shape)
                h.check(L.H5Dwrite(did, tid, msid, sel, H5P_DEFAULT, arr.ctypes.data), f"write {path}")
                return None
            finally:
                if own:
                    L.H5Tclose(tid)
                if msid != H5S_ALL:
                    L.H5Sclose(msid)
        finally:
            L.H5Tclose(ftid)
            L.H5Sclose(fsid)
            L.H5Dclose(did)

    def read(self, path, index=None):
        return self._io(path, index)

    def write(self, path, data, index=None):
        self._io(path, index, data)


def _leaves(tree, prefix=""):
    """(path, value) for every non-dict value of a nested dictionary, and the
    group paths on the way, depth first in the dictionary's order."""
    for key, value in tree.items():
        path = f"{prefix}/{key}"
        if isinstance(value, dict):
            yield path, None
            yield from _leaves(value, path)
        else:
            yield path, value


def _dataset_spec(value, n_events, per_event):
    """Shape, dtype and CXI "axes" label of the dataset that holds `value`:
    one entry per event (leading axis n_events) when `per_event`, else the
    value alone (a scalar becomes a one-element vector)."""
    arr = np.asarray(value)
    if arr.dtype.kind in "SUO":
        raise TypeError("CXI: string and object data are not supported")
    if arr.ndim == 0:
        return ((n_events,) if per_event else (1,)), arr.dtype, "experiment_identifier:value"
    shape = ((n_events,) if per_event else ()) + arr.shape
    return shape, arr.dtype, "experiment_identifier:y:x"


class CXIWriter:
    """Writes a CXI file of `N` events (reference interface:
    src/_spimage_cxi.py:5-72).

    Nested dictionaries map onto HDF5 groups.  write_stack takes arrays whose
    first axis runs over the N events and writes each one in a single call;
    write_slice writes one event's values (i given) or values that belong to
    the whole run (i None); write_to_dataset is the single-dataset form of
    both.  Datasets are created on first use with the CXI "axes" attribute."""

    def __init__(self, filename, N, logger=None):
        self.filename = os.path.expandvars(filename)
        self.N = int(N)
        self.logger = logger
        self.f = H5File(self.filename, "w")

    def _ensure(self, path, value, per_event):
        if path in self.f:
            return
        shape, dtype, axes = _dataset_spec(value, self.N, per_event)
        self.f.create_dataset(path, shape, dtype, axes=axes)
        if self.logger is not None:
            self.logger.debug("created dataset %s %s %s", path, shape, dtype)

    def write_stack(self, data, name=""):
        for path, value in _leaves(data, name):
            if value is None:
                self.f.create_group(path)
                continue
            stack = np.asarray(value)
            if stack.ndim == 0 or stack.shape[0] != self.N:
                raise ValueError(f"CXIWriter.write_stack: {path} needs {self.N} events along its first axis, "
                                 f"got shape {stack.shape}")
            self._ensure(path, stack[0], per_event=True)
            self.f.write(path, stack)

    def write_slice(self, data, name="", i=None):
        for path, value in _leaves(data, name):
            if value is None:
                self.f.create_group(path)
            else:
                self.write_to_dataset(value, path, i=i)

    def write_to_dataset(self, data, name, i=None):
        if self.logger is not None:
            self.logger.debug("writing %s (event %s)", name, i)
        self._ensure(name, data, per_event=i is not None)
        arr = np.asarray(data)
        if i is None:
            self.f.write(name, arr.reshape(1) if arr.ndim == 0 else arr)
        else:
            self.f.write(name, arr, index=int(i))

    def close(self):
        self.f.close()


class CXIReader:
    """Reads events from one .cxi file or from every .cxi file of a
    directory (reference interface: src/_spimage_cxi.py:74-292).

    The events to deliver are decided once, at construction, as index arrays
    over all files: a global keep-mask (filters on per-event datasets, then
    the first `nevents` kept events from `ifirst` on, or a random subset of
    `nevents`), and within each file the delivery order (natural, or by
    ascending `sorting_dataset`).  get_next() then walks that schedule and
    reads the stacked datasets at the scheduled event.

    Keyword arguments: nevents (0 = all), ifirst, def_stack_ds (dataset whose
    first axis counts a file's events; default "/i"), pick_events
    ("in_sequence" or "random"), event_filters ({name: {"dataset_name": ...,
    "vmin": ..., "vmax": ...} or {..., "indices": value or list}}),
    sorting_dataset, dsnames_single ({key: dataset} read whole with every
    event), logger."""

    def __init__(self, location, dsnames_stack={}, **kwargs):
        self.logger = kwargs.get("logger", None)
        self.dsnames_stack = dsnames_stack
        self.dsnames_single = kwargs.get("dsnames_single", None)
        self.def_stack_ds = kwargs.get("def_stack_ds", "/i")
        nevents = int(kwargs.get("nevents", 0))
        ifirst = int(kwargs.get("ifirst", 0))
        self.directories, self.filenames = self._resolve_location(location)
        self.Nfiles = len(self.filenames)
        self.Nevents_files = self.get_Nevents_files()
        self.Nevents_tot = int(sum(self.Nevents_files))
        if nevents < 0:
            raise ValueError(f"CXIReader: nevents must be >= 0 (got {nevents})")
        if nevents + ifirst > self.Nevents_tot:
            raise ValueError(f"CXIReader: ifirst {ifirst} + nevents {nevents} exceeds the "
                             f"{self.Nevents_tot} events found")
        starts = np.concatenate([[0], np.cumsum(self.Nevents_files)]).astype(np.int64)
        keep = self._filtered(kwargs.get("event_filters", {}), starts)
        keep = self._picked(keep, kwargs.get("pick_events", "in_sequence"), ifirst, nevents)
        self.is_event_to_process = [keep[starts[f]:starts[f + 1]] for f in range(self.Nfiles)]
        self.order = self._orders(kwargs.get("sorting_dataset", None))
        # the delivery schedule: (file, event in file) of every event to read
        files, events = [], []
        for f in range(self.Nfiles):
            sel = self.order[f][self.is_event_to_process[f][self.order[f]]]
            files.append(np.full(sel.size, f, np.int64))
            events.append(sel.astype(np.int64))
        self._sched_file = np.concatenate(files) if files else np.zeros(0, np.int64)
        self._sched_event = np.concatenate(events) if events else np.zeros(0, np.int64)
        self.Nevents_process = int(self._sched_file.size)
        self.ievent_process = -1
        self.ifile_opened = None
        self.F = None

    # -- the schedule ------------------------------------------------------
    def _path(self, f):
        return os.path.join(self.directories[f], self.filenames[f])

    def _resolve_location(self, location0):
        location = os.path.expandvars(location0)
        if os.path.isdir(location):
            names = sorted(n for n in os.listdir(location) if n.endswith(".cxi"))
            return [location] * len(names), names
        return [os.path.dirname(location)], [os.path.basename(location)]

    def get_Nevents_files(self):
        counts = []
        for f in range(len(self.filenames)):
            with H5File(self._path(f)) as h:
                counts.append(h.shape(self.def_stack_ds)[0])
        if self.logger is not None:
            for name, n in zip(self.filenames, counts):
                self.logger.info("%s: %d events", name, n)
        return counts

    def _filtered(self, event_filters, starts):
        keep = np.ones(self.Nevents_tot, bool)
        for f in range(self.Nfiles):
            part = keep[starts[f]:starts[f + 1]]
            with H5File(self._path(f)) as h:
                for fname, flt in event_filters.items():
                    values = np.ravel(h.read(flt["dataset_name"]))
                    if "vmin" in flt and "vmax" in flt:
                        ok = (values >= flt["vmin"]) & (values <= flt["vmax"])
                    elif "indices" in flt:
                        ok = np.isin(values, np.atleast_1d(np.asarray(flt["indices"])))
                    else:
                        if self.logger is not None:
                            self.logger.warning("filter %s has neither vmin/vmax nor indices", fname)
                        continue
                    part &= ok
        return keep

    def _picked(self, keep, mode, ifirst, nevents):
        if mode == "in_sequence":
            keep[:ifirst] = False
            if nevents:
                kept = np.flatnonzero(keep)
                keep[kept[nevents:]] = False
        elif mode == "random":
            kept = np.flatnonzero(keep)
            chosen = np.random.default_rng().permutation(kept)[:nevents]
            keep = np.zeros_like(keep)
            keep[chosen] = True
        else:
            raise ValueError(f"CXIReader: unknown pick_events mode {mode!r}")
        return keep

    def _orders(self, sorting_dsname):
        if sorting_dsname is None:
            return [np.arange(n) for n in self.Nevents_files]
        out = []
        for f in range(self.Nfiles):
            with H5File(self._path(f)) as h:
                out.append(np.argsort(np.ravel(h.read(sorting_dsname)), kind="stable"))
        return out

    # -- delivery ----------------------------------------------------------
    def get_next(self):
        if self.ievent_process + 1 >= self.Nevents_process:
            self.close()
            return None
        self.ievent_process += 1
        f = int(self._sched_file[self.ievent_process])
        e = int(self._sched_event[self.ievent_process])
        if self.ifile_opened != f:
            self.close()
            self.F = H5File(self._path(f))
            self.ifile_opened = f
        event = {"i": self.ievent_process, "filename": self.filenames[f], "filename_i": f}
        for key, dsname in (self.dsnames_stack or {}).items():
            event[key] = self.F.read(dsname, index=e)
        for key, dsname in (self.dsnames_single or {}).items():
            event[key] = self.F.read(dsname)
        return event

    def close(self):
        if self.F is not None:
            self.F.close()
            self.F = None
            self.ifile_opened = None
