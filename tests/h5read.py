"""A reader for HDF5 files through the HDF5 C library (ctypes): the few calls
the Yell 1.0 tests need to list datasets and read their types and values."""
import ctypes as C
import os

import numpy as np

from diffusecode_amd import build as B

ROOT = B.hdf5_root()

hid_t = C.c_int64
hsize_t = C.c_uint64
H5T_INTEGER, H5T_FLOAT, H5T_STRING = 0, 1, 3
H5T_ORDER_LE, H5T_ORDER_BE = 0, 1
H5T_STR_NULLTERM = 0
H5_VARIABLE = C.c_size_t(-1).value


class H5:
    """The few HDF5 C calls a reader needs."""

    def __init__(self):
        L = C.CDLL(os.path.join(ROOT, "lib", "libhdf5.so"))
        L.H5open()
        for f, res, args in (("H5Fopen", hid_t, [C.c_char_p, C.c_uint, hid_t]), ("H5Fclose", C.c_int, [hid_t]),
                             ("H5Dopen2", hid_t, [hid_t, C.c_char_p, hid_t]), ("H5Dclose", C.c_int, [hid_t]),
                             ("H5Dget_type", hid_t, [hid_t]), ("H5Dget_space", hid_t, [hid_t]),
                             ("H5Tget_class", C.c_int, [hid_t]), ("H5Tget_size", C.c_size_t, [hid_t]),
                             ("H5Tget_order", C.c_int, [hid_t]), ("H5Tis_variable_str", C.c_int, [hid_t]),
                             ("H5Tget_strpad", C.c_int, [hid_t]), ("H5Tclose", C.c_int, [hid_t]),
                             ("H5Tcopy", hid_t, [hid_t]), ("H5Tset_size", C.c_int, [hid_t, C.c_size_t]),
                             ("H5Sget_simple_extent_ndims", C.c_int, [hid_t]),
                             ("H5Sget_simple_extent_dims", C.c_int, [hid_t, C.POINTER(hsize_t), C.c_void_p]),
                             ("H5Sclose", C.c_int, [hid_t]),
                             ("H5Dread", C.c_int, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
                             ("H5Lget_name_by_idx", C.c_ssize_t,
                              [hid_t, C.c_char_p, C.c_int, C.c_int, hsize_t, C.c_char_p, C.c_size_t, hid_t]),
                             ("H5Lexists", C.c_int, [hid_t, C.c_char_p, hid_t])):
            fn = getattr(L, f)
            fn.restype, fn.argtypes = res, args
        self.L = L
        g = lambda n: hid_t.in_dll(L, n).value
        self.NATIVE_DOUBLE, self.NATIVE_SHORT, self.C_S1 = g("H5T_NATIVE_DOUBLE_g"), g("H5T_NATIVE_SHORT_g"), g("H5T_C_S1_g")

    def names(self, f):
        out, i = [], 0
        while True:
            buf = C.create_string_buffer(256)
            n = self.L.H5Lget_name_by_idx(f, b".", 0, 0, i, buf, 256, 0)
            if n < 0:
                return set(out)
            out.append(buf.value.decode())
            i += 1

    def info(self, f, name):
        d = self.L.H5Dopen2(f, name.encode(), 0)
        assert d >= 0, name
        t, s = self.L.H5Dget_type(d), self.L.H5Dget_space(d)
        nd = self.L.H5Sget_simple_extent_ndims(s)
        dims = (hsize_t * max(nd, 1))()
        self.L.H5Sget_simple_extent_dims(s, dims, None)
        cls = self.L.H5Tget_class(t)
        r = dict(cls=cls, size=self.L.H5Tget_size(t), dims=tuple(dims[:nd]),
                 order=self.L.H5Tget_order(t) if cls != H5T_STRING else None,
                 vlen=bool(self.L.H5Tis_variable_str(t)) if cls == H5T_STRING else None,
                 strpad=self.L.H5Tget_strpad(t) if cls == H5T_STRING else None)
        self.L.H5Tclose(t)
        self.L.H5Sclose(s)
        return d, r

    def doubles(self, f, name):
        d, r = self.info(f, name)
        out = np.zeros(int(np.prod(r["dims"])) if r["dims"] else 1)
        assert self.L.H5Dread(d, self.NATIVE_DOUBLE, 0, 0, 0, out.ctypes.data) >= 0
        self.L.H5Dclose(d)
        return out.reshape(r["dims"]) if r["dims"] else out, r

    def short(self, f, name):
        d, r = self.info(f, name)
        out = np.zeros(1, np.int16)
        assert self.L.H5Dread(d, self.NATIVE_SHORT, 0, 0, 0, out.ctypes.data) >= 0
        self.L.H5Dclose(d)
        return int(out[0]), r

    def vlstring(self, f, name):
        d, r = self.info(f, name)
        mt = self.L.H5Tcopy(self.C_S1)
        self.L.H5Tset_size(mt, H5_VARIABLE)
        p = C.c_char_p()
        assert self.L.H5Dread(d, mt, 0, 0, 0, C.addressof(p)) >= 0
        self.L.H5Tclose(mt)
        self.L.H5Dclose(d)
        return p.value.decode(), r

    def fixstring(self, f, name):
        d, r = self.info(f, name)
        buf = C.create_string_buffer(r["size"] * int(np.prod(r["dims"])))
        mt = self.L.H5Tcopy(self.C_S1)
        self.L.H5Tset_size(mt, r["size"])
        assert self.L.H5Dread(d, mt, 0, 0, 0, buf) >= 0
        self.L.H5Tclose(mt)
        self.L.H5Dclose(d)
        return buf.raw, r

    def open(self, path):
        f = self.L.H5Fopen(os.fsencode(path), 0, 0)
        assert f >= 0
        return f
