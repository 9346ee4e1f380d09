"""Minimal HDF5 file access over the C library (libhdf5 >= 1.10) via ctypes.

The snapshot writer of the facade needs what abie/data_io.py takes from
h5py: create a file, a float attribute on the root group, groups, and
contiguous little-endian float64 / int64 datasets written in one piece; and
for reading back (tests, ``snapshot_serialization``) the same objects.
h5py is not installed on this image, the HDF5 C library is: this module
binds the handful of C entry points that cover those operations.

The library is looked up in ``$ABIE_HDF5_LIB``, then by ``find_library``,
then at a few fixed install locations; ``available()`` says whether one was
found.  Every failing call raises ``H5Error`` (the library's own error
printing is switched off).
"""
import ctypes
import ctypes.util
import os

import numpy as np

_CANDIDATES = (
    "/opt/conda/lib/libhdf5.so",
    "/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5.so",
    "/usr/lib/x86_64-linux-gnu/libhdf5_serial.so",
    "/usr/lib64/libhdf5.so",
    "/usr/local/lib/libhdf5.so",
)

# Constants of H5Fpublic.h / H5Ppublic.h / H5Spublic.h / H5Tpublic.h /
# H5public.h (values fixed by the 1.10 ABI).
_H5F_ACC_RDONLY = 0x0000
_H5F_ACC_RDWR = 0x0001
_H5F_ACC_TRUNC = 0x0002
_H5P_DEFAULT = 0
_H5S_ALL = 0
_H5S_SCALAR = 0
_H5F_SCOPE_LOCAL = 0
_H5T_INTEGER = 0
_H5T_FLOAT = 1
_H5_INDEX_NAME = 0
_H5_ITER_INC = 0


class H5Error(RuntimeError):
    pass


class _Lib(object):
    def __init__(self, path):
        lib = ctypes.CDLL(path)
        if lib.H5open() < 0:
            raise H5Error("H5open failed in %s" % path)
        maj, mn, rel = ctypes.c_uint(), ctypes.c_uint(), ctypes.c_uint()
        lib.H5get_libversion(ctypes.byref(maj), ctypes.byref(mn), ctypes.byref(rel))
        self.version = (maj.value, mn.value, rel.value)
        if self.version < (1, 10, 0):
            raise H5Error("libhdf5 %d.%d.%d at %s: 1.10 or newer is needed (64-bit hid_t)"
                          % (self.version + (path,)))
        self.path = path
        hid = ctypes.c_int64
        herr = ctypes.c_int
        hsize_p = ctypes.POINTER(ctypes.c_ulonglong)
        cs = ctypes.c_char_p
        vp = ctypes.c_void_p
        sig = {
            "H5Eset_auto2": (herr, [hid, vp, vp]),
            "H5Fcreate": (hid, [cs, ctypes.c_uint, hid, hid]),
            "H5Fopen": (hid, [cs, ctypes.c_uint, hid]),
            "H5Fflush": (herr, [hid, ctypes.c_int]),
            "H5Fclose": (herr, [hid]),
            "H5Gcreate2": (hid, [hid, cs, hid, hid, hid]),
            "H5Gopen2": (hid, [hid, cs, hid]),
            "H5Gclose": (herr, [hid]),
            "H5Screate": (hid, [ctypes.c_int]),
            "H5Screate_simple": (hid, [ctypes.c_int, hsize_p, hsize_p]),
            "H5Sget_simple_extent_ndims": (ctypes.c_int, [hid]),
            "H5Sget_simple_extent_dims": (ctypes.c_int, [hid, hsize_p, hsize_p]),
            "H5Sclose": (herr, [hid]),
            "H5Dcreate2": (hid, [hid, cs, hid, hid, hid, hid, hid]),
            "H5Dopen2": (hid, [hid, cs, hid]),
            "H5Dwrite": (herr, [hid, hid, hid, hid, hid, vp]),
            "H5Dread": (herr, [hid, hid, hid, hid, hid, vp]),
            "H5Dget_space": (hid, [hid]),
            "H5Dget_type": (hid, [hid]),
            "H5Dclose": (herr, [hid]),
            "H5Tget_class": (ctypes.c_int, [hid]),
            "H5Tget_size": (ctypes.c_size_t, [hid]),
            "H5Tclose": (herr, [hid]),
            "H5Acreate2": (hid, [hid, cs, hid, hid, hid, hid]),
            "H5Aopen": (hid, [hid, cs, hid]),
            "H5Awrite": (herr, [hid, hid, vp]),
            "H5Aread": (herr, [hid, hid, vp]),
            "H5Aexists": (ctypes.c_int, [hid, cs]),
            "H5Aget_type": (hid, [hid]),
            "H5Aclose": (herr, [hid]),
            "H5Lexists": (ctypes.c_int, [hid, cs, hid]),
            "H5Literate": (herr, [hid, ctypes.c_int, ctypes.c_int, hsize_p, vp, vp]),
            "H5Oopen": (hid, [hid, cs, hid]),
            "H5Oclose": (herr, [hid]),
            "H5Iget_type": (ctypes.c_int, [hid]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        self.lib = lib
        self.NATIVE_DOUBLE = hid.in_dll(lib, "H5T_NATIVE_DOUBLE_g").value
        self.NATIVE_INT64 = hid.in_dll(lib, "H5T_NATIVE_INT64_g").value
        self.IEEE_F64LE = hid.in_dll(lib, "H5T_IEEE_F64LE_g").value
        self.STD_I64LE = hid.in_dll(lib, "H5T_STD_I64LE_g").value
        lib.H5Eset_auto2(0, None, None)

    def check(self, ret, what):
        if ret < 0:
            raise H5Error("%s failed" % what)
        return ret


_LIB = None
_LIB_ERR = None


def _find():
    env = os.environ.get("ABIE_HDF5_LIB")
    if env:
        return [env]
    out = []
    found = ctypes.util.find_library("hdf5")
    if found:
        out.append(found)
    out.extend(p for p in _CANDIDATES if os.path.exists(p))
    return out


def lib():
    """The loaded library; raises H5Error when no usable libhdf5 exists."""
    global _LIB, _LIB_ERR
    if _LIB is None and _LIB_ERR is None:
        errs = []
        for path in _find():
            try:
                _LIB = _Lib(path)
                break
            except (OSError, AttributeError, H5Error) as e:
                errs.append("%s: %s" % (path, e))
        if _LIB is None:
            _LIB_ERR = "no usable libhdf5 (set ABIE_HDF5_LIB); tried: %s" % ("; ".join(errs) or "nothing found")
    if _LIB is None:
        raise H5Error(_LIB_ERR)
    return _LIB


def available():
    try:
        lib()
        return True
    except H5Error:
        return False


def _dims(shape):
    arr = (ctypes.c_ulonglong * max(len(shape), 1))(*shape)
    return arr


class File(object):
    """An HDF5 file: mode "w" creates (truncating), "r" reads, "a" appends."""

    def __init__(self, path, mode="r"):
        self._h = lib()
        L = self._h.lib
        p = os.fsencode(path)
        if mode == "w":
            fid = L.H5Fcreate(p, _H5F_ACC_TRUNC, _H5P_DEFAULT, _H5P_DEFAULT)
        elif mode == "r":
            fid = L.H5Fopen(p, _H5F_ACC_RDONLY, _H5P_DEFAULT)
        elif mode == "a":
            fid = L.H5Fopen(p, _H5F_ACC_RDWR, _H5P_DEFAULT) if os.path.exists(path) else \
                L.H5Fcreate(p, _H5F_ACC_TRUNC, _H5P_DEFAULT, _H5P_DEFAULT)
        else:
            raise ValueError("mode must be 'r', 'w' or 'a'")
        self._fid = self._h.check(fid, "opening %s" % path)
        self.filename = path

    # ---------------------------------------------------------------- misc
    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def close(self):
        if self._fid is not None:
            self._h.check(self._h.lib.H5Fclose(self._fid), "H5Fclose")
            self._fid = None

    def flush(self):
        self._h.check(self._h.lib.H5Fflush(self._fid, _H5F_SCOPE_LOCAL), "H5Fflush")

    def __contains__(self, path):
        L = self._h.lib
        parts = [q for q in path.strip("/").split("/") if q]
        cur = ""
        for q in parts:
            cur += "/" + q
            if L.H5Lexists(self._fid, cur.encode(), _H5P_DEFAULT) <= 0:
                return False
        return True

    # --------------------------------------------------------- attributes
    def set_attr(self, name, value, obj="/"):
        """Scalar float64 or int64 attribute on a group or dataset."""
        L = self._h.lib
        value = np.asarray(value)
        if value.ndim != 0:
            raise ValueError("only scalar attributes are supported")
        if np.issubdtype(value.dtype, np.integer):
            ftype, mtype, buf = self._h.STD_I64LE, self._h.NATIVE_INT64, np.array(value, dtype=np.int64)
        else:
            ftype, mtype, buf = self._h.IEEE_F64LE, self._h.NATIVE_DOUBLE, np.array(value, dtype=np.float64)
        oid = self._h.check(L.H5Oopen(self._fid, obj.encode(), _H5P_DEFAULT), "H5Oopen %s" % obj)
        try:
            if L.H5Aexists(oid, name.encode()) > 0:
                raise H5Error("attribute %s exists on %s" % (name, obj))
            sid = self._h.check(L.H5Screate(_H5S_SCALAR), "H5Screate")
            try:
                aid = self._h.check(L.H5Acreate2(oid, name.encode(), ftype, sid, _H5P_DEFAULT, _H5P_DEFAULT),
                                    "H5Acreate2 %s" % name)
                try:
                    self._h.check(L.H5Awrite(aid, mtype, buf.ctypes.data), "H5Awrite %s" % name)
                finally:
                    L.H5Aclose(aid)
            finally:
                L.H5Sclose(sid)
        finally:
            L.H5Oclose(oid)

    def get_attr(self, name, obj="/"):
        L = self._h.lib
        oid = self._h.check(L.H5Oopen(self._fid, obj.encode(), _H5P_DEFAULT), "H5Oopen %s" % obj)
        try:
            aid = self._h.check(L.H5Aopen(oid, name.encode(), _H5P_DEFAULT), "H5Aopen %s" % name)
            try:
                # the attribute's class decides the memory type
                out = np.zeros((), dtype=np.float64)
                if self._attr_is_int(aid):
                    out = np.zeros((), dtype=np.int64)
                    self._h.check(L.H5Aread(aid, self._h.NATIVE_INT64, out.ctypes.data), "H5Aread")
                else:
                    self._h.check(L.H5Aread(aid, self._h.NATIVE_DOUBLE, out.ctypes.data), "H5Aread")
                return out[()]
            finally:
                L.H5Aclose(aid)
        finally:
            L.H5Oclose(oid)

    def _attr_is_int(self, aid):
        L = self._h.lib
        tid = self._h.check(L.H5Aget_type(aid), "H5Aget_type")
        try:
            return L.H5Tget_class(tid) == _H5T_INTEGER
        finally:
            L.H5Tclose(tid)

    # ------------------------------------------------------ groups, data
    def create_group(self, path):
        L = self._h.lib
        gid = self._h.check(L.H5Gcreate2(self._fid, path.encode(), _H5P_DEFAULT, _H5P_DEFAULT, _H5P_DEFAULT),
                            "H5Gcreate2 %s" % path)
        L.H5Gclose(gid)

    def create_dataset(self, path, data):
        """Write ``data`` (float or integer array, any rank) as a contiguous
        little-endian float64 / int64 dataset at ``path``."""
        L = self._h.lib
        a = np.asarray(data)
        if np.issubdtype(a.dtype, np.integer) or a.dtype == np.bool_:
            a = np.ascontiguousarray(a, dtype=np.int64)
            ftype, mtype = self._h.STD_I64LE, self._h.NATIVE_INT64
        else:
            a = np.ascontiguousarray(a, dtype=np.float64)
            ftype, mtype = self._h.IEEE_F64LE, self._h.NATIVE_DOUBLE
        if a.ndim == 0:
            sid = self._h.check(L.H5Screate(_H5S_SCALAR), "H5Screate")
        else:
            d = _dims(a.shape)
            sid = self._h.check(L.H5Screate_simple(a.ndim, d, d), "H5Screate_simple")
        try:
            did = self._h.check(L.H5Dcreate2(self._fid, path.encode(), ftype, sid, _H5P_DEFAULT, _H5P_DEFAULT,
                                             _H5P_DEFAULT), "H5Dcreate2 %s" % path)
            try:
                if a.size:
                    self._h.check(L.H5Dwrite(did, mtype, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, a.ctypes.data),
                                  "H5Dwrite %s" % path)
            finally:
                L.H5Dclose(did)
        finally:
            L.H5Sclose(sid)

    def read(self, path):
        """The dataset at ``path`` as a numpy array (float64 or int64)."""
        L = self._h.lib
        did = self._h.check(L.H5Dopen2(self._fid, path.encode(), _H5P_DEFAULT), "H5Dopen2 %s" % path)
        try:
            tid = self._h.check(L.H5Dget_type(did), "H5Dget_type")
            try:
                cls = L.H5Tget_class(tid)
            finally:
                L.H5Tclose(tid)
            if cls == _H5T_FLOAT:
                dtype, mtype = np.float64, self._h.NATIVE_DOUBLE
            elif cls == _H5T_INTEGER:
                dtype, mtype = np.int64, self._h.NATIVE_INT64
            else:
                raise H5Error("%s: only float and integer datasets are supported" % path)
            sid = self._h.check(L.H5Dget_space(did), "H5Dget_space")
            try:
                nd = self._h.check(L.H5Sget_simple_extent_ndims(sid), "ndims")
                d = _dims([0] * nd)
                if nd:
                    self._h.check(L.H5Sget_simple_extent_dims(sid, d, None), "dims")
                shape = tuple(int(d[k]) for k in range(nd))
            finally:
                L.H5Sclose(sid)
            out = np.empty(shape, dtype=dtype)
            if out.size:
                self._h.check(L.H5Dread(did, mtype, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, out.ctypes.data),
                              "H5Dread %s" % path)
            return out
        finally:
            L.H5Dclose(did)

    def keys(self, group="/"):
        """Names of the links in ``group``, in name order."""
        L = self._h.lib
        names = []
        cb_t = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_int64, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p)

        def cb(_gid, name, _info, _op):
            names.append(name.decode())
            return 0

        fn = cb_t(cb)
        gid = self._h.check(L.H5Gopen2(self._fid, group.encode(), _H5P_DEFAULT), "H5Gopen2 %s" % group)
        try:
            idx = ctypes.c_ulonglong(0)
            self._h.check(L.H5Literate(gid, _H5_INDEX_NAME, _H5_ITER_INC, ctypes.byref(idx),
                                       ctypes.cast(fn, ctypes.c_void_p), None), "H5Literate %s" % group)
        finally:
            L.H5Gclose(gid)
        return names

    def is_group(self, path):
        L = self._h.lib
        oid = self._h.check(L.H5Oopen(self._fid, path.encode(), _H5P_DEFAULT), "H5Oopen %s" % path)
        try:
            # H5I_GROUP = 2 in H5Ipublic.h
            return L.H5Iget_type(oid) == 2
        finally:
            L.H5Oclose(oid)
