"""CXI (HDF5) reading and writing for patterns in and reconstructions out.

Mirrors the reference's ``spimage.CXIWriter`` and ``spimage.CXIReader``
(src/_spimage_cxi.py:5-292) with the same constructors, methods and event
dictionaries.  The reference builds on h5py, which this image does not
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


class CXIWriter:
    """reference src/_spimage_cxi.py:5-72: datasets of N events, written
    whole or one event at a time; nested dictionaries become groups."""

    def __init__(self, filename, N, logger=None):
        self.filename = os.path.expandvars(filename)
        self.f = H5File(self.filename, "w")
        self.N = N
        self.logger = logger

    def write_stack(self, data, name=""):
        for i in range(self.N):
            for k, d in data.items():
                name_k = name + "/" + k
                if isinstance(d, dict):
                    if name_k not in self.f:
                        self.f.create_group(name_k)
                        self.write_stack(d, name_k)
                else:
                    d_s = d[i, :] if len(np.shape(d)) > 1 else d[i]
                    self.write_to_dataset(d_s, name_k, i=i)

    def write_slice(self, data, name="", i=None):
        for k, d in data.items():
            name_k = name + "/" + k
            if isinstance(d, dict):
                if name_k not in self.f:
                    self.f.create_group(name_k)
                self.write_slice(d, name_k, i=i)
            else:
                self.write_to_dataset(d, name_k, i=i)

    def write_to_dataset(self, data, name, i=None):
        if self.logger is not None:
            self.logger.debug("Write dataset %s of event %s." % (name, i))
        scalar = np.isscalar(data)
        if name not in self.f:
            if scalar:
                s = [1] if i is None else [self.N]
                t = np.asarray(data).dtype
                axes = "experiment_identifier:value"
            else:
                data = np.asarray(data)
                s = list(data.shape)
                if i is not None:
                    s.insert(0, self.N)
                t = data.dtype
                axes = "experiment_identifier:y:x"
            if t.kind in "SUO":
                raise TypeError(f"CXIWriter: string data ({name}) is not supported")
            self.f.create_dataset(name, tuple(s), t, axes=axes)
        if i is None:
            if scalar:
                self.f.write(name, np.asarray(data).reshape(1), index=None)
            else:
                self.f.write(name, np.asarray(data))
        else:
            self.f.write(name, np.asarray(data), index=i)

    def close(self):
        self.f.close()


class CXIReader:
    """reference src/_spimage_cxi.py:74-292: events of one .cxi file or of
    every .cxi file in a directory, with the picking ("in_sequence",
    "random"), filter (vmin/vmax, indices) and sorting options."""

    def __init__(self, location, dsnames_stack={}, **kwargs):
        self.logger = kwargs.get("logger", None)
        nevents = kwargs.get("nevents", 0)
        ifirst = kwargs.get("ifirst", 0)
        self.def_stack_ds = kwargs.get("def_stack_ds", "/i")
        pick_events = kwargs.get("pick_events", "in_sequence")
        event_filters = kwargs.get("event_filters", {})
        sorting_dsname = kwargs.get("sorting_dataset", None)
        self.dsnames_single = kwargs.get("dsnames_single", None)
        self.dsnames_stack = dsnames_stack
        self.directories, self.filenames = self._resolve_location(location)
        self.ifile = 0
        self.ifile_opened = None
        self.F = None
        self.Nfiles = len(self.filenames)
        self.Nevents_files = self.get_Nevents_files()
        self.Nevents_tot = int(sum(self.Nevents_files))
        if nevents < 0:
            sys.exit("ERROR: Events to read smaller 0. Change your configuration.")
        elif nevents + ifirst > self.Nevents_tot:
            sys.exit("ERROR: Not enough events to read. Change your configuration.")
        to_process = [np.ones(N, dtype=bool) for N in self.Nevents_files]
        to_process = self._filter_events(to_process, event_filters)
        to_process = self._pick_events(to_process, pick_events, ifirst, nevents)
        self.is_event_to_process = to_process
        self.ievent_file = -1
        self.ievent_process = -1
        self.Nevents_process = int(sum(t.sum() for t in to_process))
        self.order = self._set_order(to_process, sorting_dsname)

    def _path(self, i):
        return os.path.join(self.directories[i], self.filenames[i])

    def get_next(self):
        if self._next():
            return self._get(self.dsnames_stack, self.dsnames_single)
        return None

    def get_Nevents_files(self):
        out = []
        for i in range(len(self.filenames)):
            with H5File(self._path(i)) as f:
                out.append(f.shape(self.def_stack_ds)[0])
        return out

    def close(self):
        if self.F is not None:
            self.F.close()
            self.F = None

    def _resolve_location(self, location0):
        location = os.path.expandvars(location0)
        if os.path.isdir(location):
            fs = sorted(x for x in os.listdir(location) if x[-4:] == ".cxi")
            return [location + "/"] * len(fs), fs
        fname = location.split("/")[-1]
        return [location[:-len(fname)]], [fname]

    def _pick_events(self, to_process, mode, ifirst, nevents):
        temp = np.concatenate(to_process) if to_process else np.zeros(0, bool)
        if mode == "in_sequence":
            temp[:ifirst] = False
            if nevents != 0 and nevents < temp.sum():
                s = 0
                i = ifirst
                for i in range(ifirst, self.Nevents_tot):
                    if temp[i]:
                        s += 1
                    if s == nevents:
                        break
                temp[i + 1:] = False
        elif mode == "random":
            rng = np.random.default_rng()
            pool = list(np.arange(self.Nevents_tot)[temp])
            temp = np.zeros_like(temp)
            for _ in range(nevents):
                if len(pool) != 1:
                    temp[pool.pop(int(rng.integers(len(pool))))] = True
                else:
                    temp[pool[0]] = True
        else:
            raise ValueError(f"no valid picking mode chosen: {mode}")
        out, i = [], 0
        for N in self.Nevents_files:
            out.append(temp[i:i + N])
            i += N
        return out

    def _filter_events(self, to_process, event_filters):
        for i in range(self.Nfiles):
            with H5File(self._path(i)) as f:
                for flt in event_filters.values():
                    ds = f.read(flt["dataset_name"]).flatten()
                    if "vmin" in flt and "vmax" in flt:
                        F = (ds >= flt["vmin"]) & (ds <= flt["vmax"])
                    elif "indices" in flt:
                        F = np.zeros_like(to_process[i])
                        for index in (flt["indices"] if isinstance(flt["indices"], list) else [flt["indices"]]):
                            F[ds == index] = True
                    else:
                        if self.logger is not None:
                            self.logger.warning("No valid filter arguments given!")
                        F = np.ones_like(to_process[i])
                    to_process[i] = to_process[i] & F
        return to_process

    def _set_order(self, to_process, sorting_dsname):
        order = []
        for i, N in enumerate(self.Nevents_files):
            if sorting_dsname is None:
                order.append(np.arange(N))
            else:
                with H5File(self._path(i)) as f:
                    order.append(np.argsort(f.read(sorting_dsname)))
        return order

    def _next(self):
        while True:
            if self.ievent_process == self.Nevents_process - 1:
                return False
            self.ievent_file += 1
            if self.ifile >= self.Nfiles:
                self.close()
                return False
            if self.ievent_file >= self.Nevents_files[self.ifile]:
                self.ifile += 1
                self.ievent_file = 0
                if self.ifile >= self.Nfiles:
                    self.close()
                    return False
            if self.is_event_to_process[self.ifile][self.order[self.ifile][self.ievent_file]]:
                self.ievent_process += 1
                return True

    def _open_file(self):
        if self.ifile_opened != self.ifile:
            self.close()
            self.F = H5File(self._path(self.ifile))
            self.ifile_opened = self.ifile

    def _get(self, dsnames_stack, dsnames_single):
        self._open_file()
        D = {"i": self.ievent_process, "filename": self.filenames[self.ifile_opened],
             "filename_i": self.ifile_opened}
        if dsnames_stack is not None:
            for key, dsname in dsnames_stack.items():
                D[key] = self.F.read(dsname, index=int(self.order[self.ifile][self.ievent_file]))
        if dsnames_single is not None:
            for key, dsname in dsnames_single.items():
                D[key] = self.F.read(dsname)
        return D
