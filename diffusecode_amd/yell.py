"""DISCUS "Yell 1.0" HDF5 output through libdiffuse_yell.so (dc_write_yell,
include/diffuse_yell.h): the writer DISCUS's output menu calls for
``outfile type hdf5`` (discus/prog/graphic.f90:606-636 ->
lib_f90/hdf_write.f90:12-326).

The library is built next to this file by ``diffusecode_amd.build.build_yell``
when an HDF5 installation is present; without it the writer is absent and
``write_yell`` raises, as DISCUS without HDF5 links hdf_write_no.f90.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdiffuse_yell.so")

DC_YELL_OK = 0
DC_YELL_ERR_IO = -2
DC_YELL_ERR_ARG = -1000
VAL_PDF, VAL_3DPDF = 14, 15          # qval.f90:16-17: the direct-space values (is_direct = 1)

_lib = None


class YellDesc(C.Structure):
    _fields_ = [("inc", C.c_int * 3), ("eck", C.c_double * 12), ("vi", C.c_double * 9), ("extr", C.c_int * 3),
                ("a0", C.c_double * 3), ("win", C.c_double * 3), ("value", C.c_int), ("valmax", C.c_double)]


class YellError(RuntimeError):
    def __init__(self, status, path):
        self.status = status
        what = "HDF5 could not create or write the file" if status == DC_YELL_ERR_IO else "invalid argument"
        super().__init__(f"dc_write_yell({path!r}): {what} (status {status})")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with diffusecode_amd.build.build_yell() "
                               "(needs an HDF5 installation)")
        L = C.CDLL(LIB_PATH)
        L.dc_write_yell.restype = C.c_int
        L.dc_write_yell.argtypes = [C.c_char_p, C.POINTER(YellDesc), C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def _mat(x, n, m):
    a = np.zeros((n, m))
    x = np.asarray(x, np.float64)
    a[:x.shape[0], :x.shape[1]] = x
    return a.ravel(order="F")


def write_yell(path, qvalues, eck, vi, a0, win, value=1, valmax=0.0, extr=(1, 2, 3), sigma=None) -> None:
    """gen_hdf5_write: qvalues (inc1, inc2, inc3) per-pixel values (already
    multiplied by the output scale), eck (3, 3 or 4) corners with eck[:, 0] the
    lower-left, vi (3, 3) increments with vi[:, 0] along the abscissa, cell
    lengths a0 and angles win, the qval code ``value`` (14 PDF, 15 3DPDF mark
    the file direct-space), valmax > 0 to rescale, extr = which of h, k, l
    (1-based) runs along abscissa, ordinate, top."""
    q = np.asarray(qvalues, np.float64)
    if q.ndim != 3:
        q = q.reshape(q.shape + (1,) * (3 - q.ndim))
    inc = q.shape
    d = YellDesc()
    d.inc[:] = [int(x) for x in inc]
    d.eck[:] = _mat(eck, 3, 4).tolist()
    d.vi[:] = _mat(vi, 3, 3).tolist()
    d.extr[:] = [int(x) for x in extr]
    d.a0[:] = [float(x) for x in a0]
    d.win[:] = [float(x) for x in win]
    d.value = int(value)
    d.valmax = float(valmax)
    qf = np.asfortranarray(q)
    sf = None
    if sigma is not None:
        sf = np.asfortranarray(np.asarray(sigma, np.float64).reshape(inc, order="F"))
    st = lib().dc_write_yell(os.fsencode(path), C.byref(d), qf.ctypes.data,
                             None if sf is None else sf.ctypes.data)
    if st != DC_YELL_OK:
        raise YellError(st, path)
