"""ctypes binding of libdiffuse_turbo.so (include/diffuse_turbo.h).

The shared library is built in-tree (diffusecode_amd/build.py).  Importing
this module never falls back to anything else: if the library is missing or
cannot be loaded, ``lib()`` raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdiffuse_turbo.so")

DC_OK = 0
DC_WARN_NOCELL = 1
DC_ERR_ISTL = -3
DC_ERR_CANCEL = -14
DC_ERR_ZEROVEC = -132
DC_ERR_LOTS = -162
DC_ERR_NOATOMS = -198
DC_ERR_ARG = -1000
DC_ERR_DEVICE = -1001
DC_ERR_NOMEM = -1002
DC_ERR_NODEV = -1003
DC_ERR_STATE = -1004

PREC = {"fp32": 0, "bf16x3": 1, "fp16x2": 2, "fp16x3": 3}
LOT = {"off": 1, "box": 2, "eli": 3}
CFPKT = 9999

EXPORTS = [
    "dc_create", "dc_destroy", "dc_last_error", "dc_last_error_type", "dc_set_precision",
    "dc_get_precision", "dc_set_gemm_tile", "dc_set_slice_atoms", "dc_set_phase_dither", "dc_set_adjoint_walk", "dc_gemm_num_tiles", "dc_gemm_tile_shape", "dc_set_cancel_flag", "dc_version", "dc_device_info", "dc_measure_peaks", "dc_strucf", "dc_strucf_dev_noalloc", "dc_four_stltab",
    "dc_crystal_set", "dc_four_run", "dc_four_run_dev", "dc_last_stats", "dc_synchronize",
    "dc_model_intensity", "dc_model_vjp", "dc_model_chi2", "dc_model_jvp",
    "dc_batch_chi2", "dc_batch_chi2_dev", "dc_set_autotune", "dc_set_graphs", "dc_batch_info",
    "dc_four_post", "dc_four_post_dev", "dc_four_fill", "dc_four_fill_dev", "dc_four_accumulate_dev",
    "dc_qval", "dc_qval_dev", "dc_nufft1", "dc_nufft1_dev", "dc_four_strucf_nufft", "dc_four_strucf_nufft_dev",
    "dc_four_run_nufft", "dc_four_run_nufft_dev",
]

DC_ERR_NOSUM = -175
ACCUM = {"init": -1, "single": 0, "accum": 1, "finished": 2}
QVAL = {"inten": 1, "ampli": 2, "phase": 3, "real": 4, "imag": 5, "sq": 7, "f2aver": 9, "faver2": 10,
        "faver": 11, "iq": 13, "f2averb": 17, "faver2b": 18, "faverb": 19}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class CrystalDesc(C.Structure):
    _fields_ = [("cr_pos", _dp), ("cr_iscat", _ip), ("natoms", C.c_int64), ("icc", C.c_int * 3),
                ("ncatoms", C.c_int), ("nscat", C.c_int), ("dim0", C.c_double * 3), ("dim_lo", C.c_double * 3)]


class RunDesc(C.Structure):
    _fields_ = [("eck", C.c_double * 12), ("vi", C.c_double * 9), ("inc", C.c_int * 3), ("rten", C.c_double * 9),
                ("cfact", _dp), ("ncfact", C.c_int), ("ilots", C.c_int), ("ls", C.c_int * 3), ("nlots", C.c_int),
                ("lbeg", _ip), ("lperiod", C.c_int), ("fave", C.c_double), ("sel_cell", C.POINTER(C.c_ubyte)),
                ("four_friedel", C.c_int), ("ano", C.c_int), ("diff_scale", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("pairs", C.c_double), ("gemm_ms", C.c_double), ("total_ms", C.c_double),
                ("reduce_ms", C.c_double), ("gemm_launches", C.c_int64), ("slices", C.c_int64),
                ("forward_ms", C.c_double), ("adjoint_ms", C.c_double), ("adjoint_walk", C.c_int64),
                ("gemm_scheme", C.c_int64)]


class Peaks(C.Structure):
    _fields_ = [("mfma_bf16_tflops", C.c_double), ("mfma_f16_tflops", C.c_double), ("mfma_f32_tflops", C.c_double),
                ("hbm_read_gbs", C.c_double), ("hbm_copy_gbs", C.c_double), ("hbm_bytes", C.c_int64), ("num_cu", C.c_int)]


class ModelDesc(C.Structure):
    _fields_ = [("pos", _dp), ("type", _ip), ("occ", _dp), ("bdw", _dp), ("natoms", C.c_int64),
                ("nscat", C.c_int), ("ncfact", C.c_int), ("cfact", _dp), ("eck", C.c_double * 12),
                ("vi", C.c_double * 9), ("inc", C.c_int * 3), ("rten", C.c_double * 9)]


class BatchDesc(C.Structure):
    _fields_ = [("nmembers", C.c_int), ("atom_beg", C.POINTER(C.c_int64)), ("pos", _dp), ("type", _ip),
                ("occ", _dp), ("bdw", _dp), ("nscat", C.c_int), ("ncfact", C.c_int), ("cfact", _dp),
                ("rten", _dp), ("eck", C.c_double * 12), ("vi", C.c_double * 9), ("inc", C.c_int * 3)]


class PostDesc(C.Structure):
    _fields_ = [("num", C.c_int * 3), ("eck", C.c_double * 12), ("vi", C.c_double * 9), ("rten", C.c_double * 9),
                ("vr", C.c_double), ("fave", C.c_double), ("nlots", C.c_int), ("icc", C.c_int * 3),
                ("ls", C.c_int * 3), ("lots_off", C.c_int), ("diff_scale", C.c_double),
                ("res_sigma", C.c_double * 3), ("res_tr", C.c_double * 9), ("filter", C.c_int),
                ("rscale", C.c_double), ("damp", C.c_double), ("width", C.c_int)]


class SymmDesc(C.Structure):
    _fields_ = [("num", C.c_int * 3), ("eck", C.c_double * 12), ("vi", C.c_double * 9), ("fave", C.c_double),
                ("four_symm", C.c_int), ("nops", C.c_int), ("mats", _dp)]


class QvalDesc(C.Structure):
    _fields_ = [("num", C.c_int * 3), ("eck", C.c_double * 12), ("vi", C.c_double * 9), ("rten", C.c_double * 9),
                ("value", C.c_int), ("laver", C.c_int), ("cfact", _dp), ("cfact_pure", _dp), ("ncfact", C.c_int),
                ("nscat", C.c_int), ("amount", _dp), ("n_real_atoms", C.c_double), ("u2aver", C.c_double),
                ("faver_sign", C.c_double)]


_lib = None


def lib():
    """Load the library (raises when it is absent: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m diffusecode_amd.build`")
    L = C.CDLL(LIB_PATH)
    L.dc_create.argtypes = [C.POINTER(C.c_void_p), C.c_int]
    L.dc_destroy.argtypes = [C.c_void_p]
    L.dc_destroy.restype = None
    L.dc_last_error.argtypes = [C.c_void_p]
    L.dc_last_error.restype = C.c_char_p
    L.dc_last_error_type.argtypes = [C.c_void_p]
    L.dc_set_precision.argtypes = [C.c_void_p, C.c_int]
    L.dc_get_precision.argtypes = [C.c_void_p]
    L.dc_set_gemm_tile.argtypes = [C.c_void_p, C.c_int]
    L.dc_set_slice_atoms.argtypes = [C.c_void_p, C.c_int]
    L.dc_set_phase_dither.argtypes = [C.c_void_p, C.c_int]
    L.dc_set_adjoint_walk.argtypes = [C.c_void_p, C.c_int]
    L.dc_gemm_tile_shape.argtypes = [C.c_int, _ip, _ip, _ip]
    L.dc_set_cancel_flag.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    L.dc_version.restype = C.c_char_p
    L.dc_device_info.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.POINTER(C.c_int)]
    L.dc_measure_peaks.argtypes = [C.c_void_p, C.c_int64, C.POINTER(Peaks)]
    L.dc_strucf.argtypes = [C.c_void_p, _dp, C.c_int64, C.c_int64, _dp, _dp, _dp, _dp, _ip, _ip, _ip, _dp,
                            C.c_int64, _ip, _dp, _dp, C.c_int, _dp]
    L.dc_four_stltab.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _ip, _dp, _ip]
    L.dc_strucf_dev_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, _dp, _dp, _dp, _dp, _ip, _ip,
                                        _ip, C.c_void_p, C.c_int64, C.c_void_p, _dp, _dp, C.c_int, C.c_void_p]
    L.dc_crystal_set.argtypes = [C.c_void_p, C.POINTER(CrystalDesc)]
    L.dc_four_run.argtypes = [C.c_void_p, C.POINTER(RunDesc), _dp, _dp, _dp, _dp, _dp]
    L.dc_four_run_dev.argtypes = [C.c_void_p, C.POINTER(RunDesc), C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p]
    L.dc_last_stats.argtypes = [C.c_void_p, C.POINTER(Stats)]
    L.dc_synchronize.argtypes = [C.c_void_p]
    L.dc_model_intensity.argtypes = [C.c_void_p, C.POINTER(ModelDesc), _dp, _dp]
    L.dc_model_vjp.argtypes = [C.c_void_p, C.POINTER(ModelDesc), _dp, _dp, _dp]
    L.dc_model_chi2.argtypes = [C.c_void_p, C.POINTER(ModelDesc), _dp, _dp, _dp, _dp, _dp]
    L.dc_model_jvp.argtypes = [C.c_void_p, C.POINTER(ModelDesc), _dp, C.c_int, _dp, _dp, _dp]
    L.dc_batch_chi2.argtypes = [C.c_void_p, C.POINTER(BatchDesc), _dp, _dp, _dp, _dp, _dp]
    L.dc_batch_chi2_dev.argtypes = [C.c_void_p, C.POINTER(BatchDesc), C.c_void_p, C.c_void_p, _dp, C.c_void_p,
                                    C.c_void_p]
    L.dc_set_autotune.argtypes = [C.c_void_p, C.c_int]
    L.dc_set_graphs.argtypes = [C.c_void_p, C.c_int]
    L.dc_batch_info.argtypes = [C.c_void_p, _ip, _ip, _ip, _dp]
    L.dc_four_post.argtypes = [C.c_void_p, C.POINTER(PostDesc), _dp, _dp, _dp]
    L.dc_four_post_dev.argtypes = [C.c_void_p, C.POINTER(PostDesc), C.c_void_p, C.c_void_p, C.c_void_p]
    L.dc_four_fill.argtypes = [C.c_void_p, C.POINTER(SymmDesc), _dp, _dp]
    L.dc_four_fill_dev.argtypes = [C.c_void_p, C.POINTER(SymmDesc), C.c_void_p, C.c_void_p]
    L.dc_four_accumulate_dev.argtypes = [C.c_void_p, C.c_int, C.POINTER(SymmDesc), C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p]
    L.dc_qval.argtypes = [C.c_void_p, C.POINTER(QvalDesc), _dp, _dp, _dp, _dp]
    L.dc_qval_dev.argtypes = [C.c_void_p, C.POINTER(QvalDesc), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.dc_nufft1.argtypes = [C.c_void_p, C.c_int, C.c_int64, _dp, _dp, _dp, _dp, C.c_int, C.c_double, C.c_int64,
                            C.c_int64, C.c_int64, _dp]
    L.dc_nufft1_dev.argtypes = [C.c_void_p, C.c_int, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_int, C.c_double, C.c_int64, C.c_int64, C.c_int64, C.c_void_p]
    L.dc_four_strucf_nufft.argtypes = [C.c_void_p, C.c_int, C.c_int64, _dp, _dp, _dp, C.c_double, _ip, _ip, _dp]
    L.dc_four_run_nufft.argtypes = [C.c_void_p, C.POINTER(RunDesc), _dp, _dp, _dp]
    L.dc_four_run_nufft_dev.argtypes = [C.c_void_p, C.POINTER(RunDesc), _dp, C.c_void_p, C.c_void_p]
    L.dc_four_strucf_nufft_dev.argtypes = [C.c_void_p, C.c_int, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_double, _ip, _ip, C.c_void_p]
    _lib = L
    return L


class DCError(RuntimeError):
    """A non-zero status of the library; ``ier_num``/``ier_typ`` follow DISCUS."""

    def __init__(self, code, typ, msg):
        super().__init__(f"status {code} (ier_typ {typ}): {msg}")
        self.ier_num = code
        self.ier_typ = typ


def _d(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _i(a):
    return None if a is None else a.ctypes.data_as(_ip)


def _f64(a, n=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def _i32(a, n=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.int32).ravel(order="F"))
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


class Context:
    """One library context (one HIP stream, device buffers kept between calls)."""

    def __init__(self, device: int = -1, precision: str = "fp32"):
        L = lib()
        h = C.c_void_p()
        st = L.dc_create(C.byref(h), device)
        if st != 0:
            raise DCError(st, 99, "dc_create failed (no gfx950 device?)")
        self.h = h
        self._keep = []
        self.set_precision(precision)

    def close(self):
        if getattr(self, "h", None):
            lib().dc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def check(self, st, allow_warn=True):
        if st < 0 or (st > 0 and not allow_warn):
            L = lib()
            raise DCError(st, L.dc_last_error_type(self.h), L.dc_last_error(self.h).decode())
        return st

    def set_precision(self, precision):
        p = PREC[precision] if isinstance(precision, str) else int(precision)
        self.check(lib().dc_set_precision(self.h, p))
        self.precision = precision

    def set_gemm_tile(self, tile):
        self.check(lib().dc_set_gemm_tile(self.h, int(tile)))

    def set_slice_atoms(self, n):
        """Atoms per GEMM slice: 0 = planned, n > 0 forced."""
        self.check(lib().dc_set_slice_atoms(self.h, int(n)))

    def set_phase_dither(self, on):
        """Per-atom phase rotation of the phase GEMM (default on)."""
        self.check(lib().dc_set_phase_dither(self.h, 1 if on else 0))

    def set_adjoint_walk(self, walk):
        """Column tiles per work-group of the split reverse pass on short
        grids: 0 = planned, 1..32 forced."""
        self.check(lib().dc_set_adjoint_walk(self.h, int(walk)))

    @staticmethod
    def gemm_tiles():
        L = lib()
        out = []
        for t in range(L.dc_gemm_num_tiles()):
            r, c, n = C.c_int(), C.c_int(), C.c_int()
            L.dc_gemm_tile_shape(t, C.byref(r), C.byref(c), C.byref(n))
            out.append((r.value, c.value, n.value))
        return out

    def set_cancel_flag(self, flag_array):
        """flag_array: numpy int32 array of size 1 kept alive by the caller."""
        self._keep.append(flag_array)
        ptr = None if flag_array is None else flag_array.ctypes.data_as(_ip)
        self.check(lib().dc_set_cancel_flag(self.h, ptr))

    def stats(self):
        s = Stats()
        self.check(lib().dc_last_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def device_info(self):
        buf = C.create_string_buffer(64)
        n = C.c_int()
        self.check(lib().dc_device_info(self.h, buf, 64, C.byref(n)))
        return buf.value.decode(), n.value

    def measure_peaks(self, hbm_bytes=0):
        """Ceilings measured with plain instruction streams (dc_measure_peaks):
        MFMA TFLOP/s for BF16, FP16 and FP32 chains, HBM read and copy GB/s."""
        p = Peaks()
        self.check(lib().dc_measure_peaks(self.h, int(hbm_bytes), C.byref(p)))
        return {k: getattr(p, k) for k, _ in Peaks._fields_}

    def synchronize(self):
        self.check(lib().dc_synchronize(self.h))

    # -- layer call ----------------------------------------------------------
    def strucf(self, xat, uin, vin, win, off_shift, lmn, num, fnum=None, cfact_col=None, istl=None,
               xm=None, rten=None):
        """Drop-in four_strucf: xat (nxat,3) positions; returns tcsf (num) complex128, Fortran order."""
        x = np.asfortranarray(np.asarray(xat, np.float64).reshape(-1, 3))
        n = x.shape[0]
        num = _i32(num, 3)
        fnum = num.copy() if fnum is None else _i32(fnum, 3)
        lform = cfact_col is not None
        cf = None if cfact_col is None else np.ascontiguousarray(np.asarray(cfact_col, np.complex128)).view(np.float64)
        ncf = 0 if cf is None else cf.size // 2
        ist = None if istl is None else _i32(istl)
        out = np.zeros(2 * int(np.prod(num)))
        st = lib().dc_strucf(self.h, _d(x) if n else None, n, max(n, 1), _d(_f64(uin, 3)), _d(_f64(vin, 3)),
                             _d(_f64(win, 3)), _d(_f64(off_shift, 9)), _i(_i32(lmn, 6)), _i(num), _i(fnum),
                             _d(cf), ncf, _i(ist), _d(None if xm is None else _f64(xm, 3)),
                             _d(None if rten is None else _f64(rten, 9)), int(lform), _d(out))
        self.check(st)
        return out.view(np.complex128).reshape(tuple(num), order="F")

    def strucf_dev_noalloc(self, d_xat, nxat, ld_xat, uin, vin, win, off_shift, lmn, num, fnum=None,
                           d_cfact=None, ncfact=0, d_istl=None, xm=None, rten=None, d_tcsf=None):
        """dc_strucf_dev_noalloc: the layer call on device buffers (raw device
        addresses as ints: xat (ld,3) FP64, cfact complex, istl int32, tcsf
        complex over num), allocating nothing."""
        num = _i32(num, 3)
        fnum = num.copy() if fnum is None else _i32(fnum, 3)
        self.check(lib().dc_strucf_dev_noalloc(
            self.h, d_xat, int(nxat), int(ld_xat), _d(_f64(uin, 3)), _d(_f64(vin, 3)), _d(_f64(win, 3)),
            _d(_f64(off_shift, 9)), _i(_i32(lmn, 6)), _i(num), _i(fnum), d_cfact, int(ncfact), d_istl,
            _d(None if xm is None else _f64(xm, 3)), _d(None if rten is None else _f64(rten, 9)),
            int(d_cfact is not None), d_tcsf))

    def four_stltab(self, xm, uin, vin, win, num, rten):
        """four_stltab on the device: istl (num) int32, Fortran order (raises
        DCError -3 when a bin exceeds CFPKT)."""
        num = _i32(num, 3)
        out = np.zeros(int(np.prod(num)), np.int32)
        self.check(lib().dc_four_stltab(self.h, _d(_f64(xm, 3)), _d(_f64(uin, 3)), _d(_f64(vin, 3)),
                                        _d(_f64(win, 3)), _i(num), _d(_f64(rten, 9)), _i(out)))
        return out.reshape(tuple(num), order="F")

    # -- run level -----------------------------------------------------------
    def crystal_set(self, crystal):
        pos = np.ascontiguousarray(np.asarray(crystal.pos, np.float64))
        isc = np.ascontiguousarray(np.asarray(crystal.iscat, np.int32))
        d = CrystalDesc()
        d.cr_pos = _d(pos)
        d.cr_iscat = _i(isc)
        d.natoms = pos.shape[0]
        d.icc[:] = [int(v) for v in crystal.icc]
        d.ncatoms = int(crystal.ncatoms)
        d.nscat = int(crystal.nscat)
        d.dim0[:] = [float(v) for v in crystal.dim0]
        d.dim_lo[:] = [float(v) for v in crystal.dim_lo]
        self.check(lib().dc_crystal_set(self.h, C.byref(d)))
        self.crystal = crystal

    @staticmethod
    def run_desc(eck, vi, inc, rten, cfact, ilots=1, ls=(1, 1, 1), lbeg=None, lperiod=True, fave=0.0,
                 sel_cell=None, four_friedel=True, ano=False, diff_scale=1.0):
        keep = []
        r = RunDesc()
        r.eck[:] = list(_f64(eck, 12))
        r.vi[:] = list(_f64(vi, 9))
        r.inc[:] = [int(v) for v in inc]
        r.rten[:] = list(_f64(rten, 9))
        cf = np.asfortranarray(np.asarray(cfact, np.complex128))
        cff = np.ascontiguousarray(cf.ravel(order="F")).view(np.float64)
        keep.append(cff)
        r.cfact = _d(cff)
        r.ncfact = cf.shape[0]
        r.ilots = int(ilots)
        r.ls[:] = [int(v) for v in ls]
        if lbeg is not None:
            lb = np.ascontiguousarray(np.asarray(lbeg, np.int32).reshape(-1, 3))
            keep.append(lb)
            r.lbeg = _i(lb)
            r.nlots = lb.shape[0]
        else:
            r.nlots = 1
        r.lperiod = int(bool(lperiod))
        r.fave = float(fave)
        if sel_cell is not None:
            sc = np.ascontiguousarray(np.asarray(sel_cell, np.uint8).ravel(order="F"))
            keep.append(sc)
            r.sel_cell = sc.ctypes.data_as(C.POINTER(C.c_ubyte))
        r.four_friedel = int(bool(four_friedel))
        r.ano = int(bool(ano))
        r.diff_scale = float(diff_scale)
        return r, keep

    def four_run(self, desc, want_var=False):
        num = tuple(desc.inc)
        ntot = int(np.prod(num))
        csf = np.zeros(2 * ntot)
        acsf = np.zeros(2 * ntot)
        dsi = np.zeros(ntot)
        fm = np.zeros(2 * ntot) if want_var else None
        dv = np.zeros(ntot) if want_var else None
        st = self.check(lib().dc_four_run(self.h, C.byref(desc), _d(csf), _d(acsf), _d(dsi), _d(fm), _d(dv)))
        out = {"status": st,
               "csf": csf.view(np.complex128).reshape(num, order="F"),
               "acsf": acsf.view(np.complex128).reshape(num, order="F"),
               "dsi": dsi.reshape(num, order="F")}
        if want_var:
            out["fmean"] = fm.view(np.complex128).reshape(num, order="F")
            out["dvar"] = dv.reshape(num, order="F")
        return out

    def four_run_dev(self, desc, csf=None, acsf=None, dsi=None, fmean=None, dvar=None):
        """Device-pointer variant: arguments are torch CUDA tensors (float64 /
        view_as_real complex128) or raw integer device addresses."""
        def p(t):
            if t is None:
                return None
            return t if isinstance(t, int) else t.data_ptr()
        return self.check(lib().dc_four_run_dev(self.h, C.byref(desc), p(csf), p(acsf), p(dsi), p(fmean), p(dvar)))

    # -- refinement model ----------------------------------------------------
    @staticmethod
    def model_desc(pos, types, occ, bdw, cfact, eck, vi, inc, rten):
        """dc_model_desc from arrays: pos (N,3) fractional, types (N) 1-based,
        occ (N), bdw (N) or None, cfact (ncfact, nscat) complex.  Returns the
        descriptor and the arrays it points to (keep them alive)."""
        keep = []
        m = ModelDesc()
        p = np.ascontiguousarray(np.asarray(pos, np.float64).reshape(-1, 3))
        t = np.ascontiguousarray(np.asarray(types, np.int32).ravel())
        o = np.ascontiguousarray(np.asarray(occ, np.float64).ravel())
        keep += [p, t, o]
        m.pos, m.type, m.occ = _d(p), _i(t), _d(o)
        if bdw is not None:
            b = np.ascontiguousarray(np.asarray(bdw, np.float64).ravel())
            keep.append(b)
            m.bdw = _d(b)
        m.natoms = p.shape[0]
        cf = np.asfortranarray(np.asarray(cfact, np.complex128))
        cff = np.ascontiguousarray(cf.ravel(order="F")).view(np.float64)
        keep.append(cff)
        m.cfact = _d(cff)
        m.ncfact, m.nscat = cf.shape[0], cf.shape[1]
        m.eck[:] = list(_f64(eck, 12))
        m.vi[:] = list(_f64(vi, 9))
        m.inc[:] = [int(v) for v in inc]
        m.rten[:] = list(_f64(rten, 9))
        return m, keep

    def model_intensity(self, desc):
        """(I, F) over the grid, Fortran order of inc."""
        num = tuple(desc.inc)
        n = int(np.prod(num))
        icalc = np.zeros(n)
        fcalc = np.zeros(2 * n)
        self.check(lib().dc_model_intensity(self.h, C.byref(desc), _d(icalc), _d(fcalc)))
        return icalc.reshape(num, order="F"), fcalc.view(np.complex128).reshape(num, order="F")

    def model_vjp(self, desc, fcalc, gi):
        """grad (N, 5): dL/dx, dL/dy, dL/dz, dL/docc, dL/dB for dL/dI = gi."""
        f = np.ascontiguousarray(np.asarray(fcalc, np.complex128).ravel(order="F")).view(np.float64)
        g = _f64(gi)
        grad = np.zeros(5 * desc.natoms)
        self.check(lib().dc_model_vjp(self.h, C.byref(desc), _d(f), _d(g), _d(grad)))
        return grad.reshape(-1, 5)

    def model_jvp(self, desc, fcalc, dparam, want_df=False):
        """dI (P, ndir) [and dF] along directions dparam (ndir, N, 5) (grad layout)."""
        dp = np.ascontiguousarray(np.asarray(dparam, np.float64).reshape(-1, desc.natoms, 5))
        nd = dp.shape[0]
        f = np.ascontiguousarray(np.asarray(fcalc, np.complex128).ravel(order="F")).view(np.float64)
        P = int(np.prod(list(desc.inc)))
        di = np.zeros(P * nd)
        df = np.zeros(2 * P * nd) if want_df else None
        self.check(lib().dc_model_jvp(self.h, C.byref(desc), _d(f), nd, _d(dp), _d(di), _d(df)))
        di = di.reshape((P, nd), order="F")
        if want_df:
            return di, df.view(np.complex128).reshape((P, nd), order="F")
        return di

    def model_chi2(self, desc, target, weight=None, want_grad=True):
        """(chi2, I, grad (N, 5) or None) for chi2 = sum w (I - target)^2."""
        num = tuple(desc.inc)
        t = _f64(target)
        w = None if weight is None else _f64(weight)
        icalc = np.zeros(t.size)
        grad = np.zeros(5 * desc.natoms) if want_grad else None
        chi = C.c_double()
        self.check(lib().dc_model_chi2(self.h, C.byref(desc), _d(t), _d(w), C.byref(chi), _d(icalc), _d(grad)))
        return chi.value, icalc.reshape(num, order="F"), None if grad is None else grad.reshape(-1, 5)

    # -- population scoring ----------------------------------------------------
    def set_autotune(self, on):
        self.check(lib().dc_set_autotune(self.h, int(bool(on))))

    def set_graphs(self, on):
        self.check(lib().dc_set_graphs(self.h, int(bool(on))))

    def batch_info(self):
        t, n, g, ms = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        self.check(lib().dc_batch_info(self.h, C.byref(t), C.byref(n), C.byref(g), C.byref(ms)))
        return {"tile": t.value, "nwg": abs(n.value), "queue": n.value < 0, "graph": bool(g.value),
                "tune_ms": ms.value}

    @staticmethod
    def batch_desc(atom_beg, pos, types, occ, bdw, cfact, rten, eck, vi, inc):
        """dc_batch_desc: atom_beg (M+1), pos (N,3), types/occ/bdw (N) member
        after member, cfact (ncfact, nscat, M) complex, rten (M, 3, 3)."""
        keep = []
        d = BatchDesc()
        ab = np.ascontiguousarray(np.asarray(atom_beg, np.int64))
        p = np.ascontiguousarray(np.asarray(pos, np.float64).reshape(-1, 3))
        t = np.ascontiguousarray(np.asarray(types, np.int32).ravel())
        o = np.ascontiguousarray(np.asarray(occ, np.float64).ravel())
        keep += [ab, p, t, o]
        d.nmembers = ab.size - 1
        d.atom_beg = ab.ctypes.data_as(C.POINTER(C.c_int64))
        d.pos, d.type, d.occ = _d(p), _i(t), _d(o)
        if bdw is not None:
            b = np.ascontiguousarray(np.asarray(bdw, np.float64).ravel())
            keep.append(b)
            d.bdw = _d(b)
        cf = np.asarray(cfact, np.complex128)
        cff = np.ascontiguousarray(cf.ravel(order="F")).view(np.float64)
        rt = np.ascontiguousarray(np.asarray(rten, np.float64).reshape(-1, 3, 3).transpose(0, 2, 1)).ravel()
        keep += [cff, rt]
        d.cfact, d.rten = _d(cff), _d(rt)
        d.ncfact, d.nscat = cf.shape[0], cf.shape[1]
        d.eck[:] = list(_f64(eck, 12))
        d.vi[:] = list(_f64(vi, 9))
        d.inc[:] = [int(v) for v in inc]
        return d, keep

    def batch_chi2(self, desc, target, weight=None, want_icalc=False, want_grad=True):
        """(chi2 (M,), I (P, M) or None, grad (N, 5) or None); target (P, M)
        with P in Fortran order of inc."""
        M = desc.nmembers
        P = int(np.prod(list(desc.inc)))
        t = np.asfortranarray(np.asarray(target, np.float64).reshape(P, M, order="F"))
        w = None if weight is None else np.asfortranarray(np.asarray(weight, np.float64).reshape(P, M, order="F"))
        chi = np.zeros(M)
        ic = np.zeros((P, M), order="F") if want_icalc else None
        n = int(desc.atom_beg[M])
        g = np.zeros(5 * n) if want_grad else None
        self.check(lib().dc_batch_chi2(self.h, C.byref(desc), _d(t), None if w is None else _d(w), _d(chi),
                                       None if ic is None else _d(ic), None if g is None else _d(g)))
        return chi, ic, None if g is None else g.reshape(n, 5)

    def batch_chi2_dev(self, desc, pos, occ, bdw, target, weight=None, icalc=None, grad=None):
        """Device variant: pos (N, 3), occ (N), bdw (N) or None, target and
        weight (M, P) (member-major, abscissa fastest), icalc (M, P) and grad
        (N, 5) float64 CUDA tensors or addresses; the other fields of desc
        (atom_beg, types, cfact, rten, grid) are the host ones of batch_desc.
        Returns chi2 (M,) as a host array."""
        def p(t):
            return None if t is None else (t if isinstance(t, int) else t.data_ptr())
        d = BatchDesc.from_buffer_copy(desc)
        d.pos = C.cast(C.c_void_p(p(pos)), _dp)
        d.occ = C.cast(C.c_void_p(p(occ)), _dp)
        d.bdw = C.cast(C.c_void_p(p(bdw)), _dp) if bdw is not None else None
        chi = np.zeros(desc.nmembers)
        self.check(lib().dc_batch_chi2_dev(self.h, C.byref(d), p(target), p(weight), _d(chi), p(icalc), p(grad)))
        return chi

    # -- post-processing (four_run tail, four_accumulate, qval) ---------------
    @staticmethod
    def post_desc(num, eck, vi, rten, vr, fave=0.0, nlots=1, icc=(1, 1, 1), ls=(1, 1, 1), lots_off=True,
                  diff_scale=1.0, res_sigma=(0.0, 0.0, 0.0), res_tr=None, lanczos=False, rscale=1.0, damp=0.5,
                  width=4):
        d = PostDesc()
        d.num[:] = [int(v) for v in num]
        d.eck[:] = list(_f64(eck, 12))
        d.vi[:] = list(_f64(vi, 9))
        d.rten[:] = list(_f64(rten, 9))
        d.vr = float(vr)
        d.fave = float(fave)
        d.nlots = int(nlots)
        d.icc[:] = [int(v) for v in icc]
        d.ls[:] = [int(v) for v in ls]
        d.lots_off = int(bool(lots_off))
        d.diff_scale = float(diff_scale)
        d.res_sigma[:] = [float(v) for v in res_sigma]
        d.res_tr[:] = list(_f64(np.eye(3) if res_tr is None else res_tr, 9))
        d.filter = 1 if lanczos else 0
        d.rscale, d.damp, d.width = float(rscale), float(damp), int(width)
        return d

    def four_post(self, desc, csf, dsi, want_dsi3d=True):
        """Host arrays (shape num, Fortran order); returns new (csf, dsi, dsi3d)."""
        num = tuple(desc.num)
        c = None if csf is None else np.ascontiguousarray(np.asarray(csf, np.complex128).ravel(order="F")).view(np.float64).copy()
        d = np.ascontiguousarray(np.asarray(dsi, np.float64).ravel(order="F")).copy()
        d3 = np.zeros_like(d) if want_dsi3d else None
        self.check(lib().dc_four_post(self.h, C.byref(desc), _d(c), _d(d), _d(d3)))
        out_c = None if c is None else c.view(np.complex128).reshape(num, order="F")
        return out_c, d.reshape(num, order="F"), None if d3 is None else d3.reshape(num, order="F")

    def four_post_dev(self, desc, csf, dsi, dsi3d=None):
        """Device variant: torch CUDA tensors (float64; complex as view_as_real) or addresses."""
        def p(t):
            return None if t is None else (t if isinstance(t, int) else t.data_ptr())
        return self.check(lib().dc_four_post_dev(self.h, C.byref(desc), p(csf), p(dsi), p(dsi3d)))

    @staticmethod
    def symm_desc(num, eck, vi, fave=0.0, four_symm=False, ops=None):
        """ops: list of 3x3 operations of the point-group part, identity first
        (DISCUS's spc_mat(1:3,1:3,1:spc_n/n_center)); nops = len(ops)."""
        d = SymmDesc()
        keep = []
        d.num[:] = [int(v) for v in num]
        d.eck[:] = list(_f64(eck, 12))
        d.vi[:] = list(_f64(vi, 9))
        d.fave = float(fave)
        d.four_symm = int(bool(four_symm))
        if ops is not None and len(ops) > 0:
            d.nops = len(ops)
            m = np.ascontiguousarray(np.stack([np.asarray(o, float) for o in ops[1:]] or [np.eye(3)], axis=2).ravel(order="F"))
            keep.append(m)
            d.mats = _d(m)
        else:
            d.nops = 1
        return d, keep

    def four_fill(self, desc, csf, dsi):
        num = tuple(desc.num)
        c = None if csf is None else np.ascontiguousarray(np.asarray(csf, np.complex128).ravel(order="F")).view(np.float64).copy()
        d = None if dsi is None else np.ascontiguousarray(np.asarray(dsi, np.float64).ravel(order="F")).copy()
        self.check(lib().dc_four_fill(self.h, C.byref(desc), _d(c), _d(d)))
        return (None if c is None else c.view(np.complex128).reshape(num, order="F"),
                None if d is None else d.reshape(num, order="F"))

    def four_accumulate_dev(self, mode, desc, csf, dsi, csf_sum=None, dsi_sum=None):
        def p(t):
            return None if t is None else (t if isinstance(t, int) else t.data_ptr())
        m = ACCUM[mode] if isinstance(mode, str) else int(mode)
        return self.check(lib().dc_four_accumulate_dev(self.h, m, C.byref(desc), p(csf), p(dsi), p(csf_sum), p(dsi_sum)))

    @staticmethod
    def qval_desc(num, eck, vi, rten, value, laver=False, cfact=None, cfact_pure=None, amount=None,
                  n_real_atoms=1.0, u2aver=0.0, faver_sign=1.0):
        d = QvalDesc()
        keep = []
        d.num[:] = [int(v) for v in num]
        d.eck[:] = list(_f64(eck, 12))
        d.vi[:] = list(_f64(vi, 9))
        d.rten[:] = list(_f64(rten, 9))
        d.value = QVAL[value] if isinstance(value, str) else int(value)
        d.laver = int(bool(laver))
        if cfact is not None:
            cf = np.asarray(cfact, np.complex128)
            cp = np.asarray(cfact if cfact_pure is None else cfact_pure, np.complex128)
            a = np.ascontiguousarray(cf.ravel(order="F")).view(np.float64)
            b = np.ascontiguousarray(cp.ravel(order="F")).view(np.float64)
            am = np.ascontiguousarray(np.asarray(amount, np.float64).ravel())
            keep += [a, b, am]
            d.cfact, d.cfact_pure, d.amount = _d(a), _d(b), _d(am)
            d.ncfact, d.nscat = cf.shape[0], cf.shape[1]
        d.n_real_atoms = float(n_real_atoms)
        d.u2aver = float(u2aver)
        d.faver_sign = float(faver_sign)
        return d, keep

    def qval(self, desc, csf=None, acsf=None, dsi=None):
        num = tuple(desc.num)
        def cz(x):
            return None if x is None else np.ascontiguousarray(np.asarray(x, np.complex128).ravel(order="F")).view(np.float64)
        c, a = cz(csf), cz(acsf)
        d = None if dsi is None else np.ascontiguousarray(np.asarray(dsi, np.float64).ravel(order="F"))
        out = np.zeros(int(np.prod(num)))
        self.check(lib().dc_qval(self.h, C.byref(desc), _d(c), _d(a), _d(d), _d(out)))
        return out.reshape(num, order="F")

    # -- NUFFT technique ----------------------------------------------------------
    def nufft1(self, pts, c, iflag, tol, N):
        """Type-1 NUFFT (finufft{1,2,3}d1): pts list of dim coordinate arrays,
        c complex (M,), N mode counts; returns fk of shape N (k1 fastest)."""
        dim = len(pts)
        xs = [np.ascontiguousarray(np.asarray(p, np.float64).ravel()) for p in pts]
        cc = np.ascontiguousarray(np.asarray(c, np.complex128).ravel()).view(np.float64)
        NN = [int(n) for n in N][:dim] + [1] * (3 - dim)
        fk = np.zeros(2 * int(np.prod(NN)))
        xyz = xs + [None] * (3 - dim)
        self.check(lib().dc_nufft1(self.h, dim, len(xs[0]), _d(xyz[0]), _d(xyz[1]), _d(xyz[2]), _d(cc), int(iflag),
                                   float(tol), NN[0], NN[1], NN[2], _d(fk)))
        return fk.view(np.complex128).reshape(tuple(NN[:dim]), order="F")

    def four_strucf_nufft(self, pts, occ, inc, lscales):
        dim = len(pts)
        xs = [np.ascontiguousarray(np.asarray(p, np.float64).ravel()) for p in pts]
        xyz = xs + [None] * (3 - dim)
        ii = np.ascontiguousarray(np.array(list(inc)[:dim] + [1] * (3 - dim), np.int32))
        ls = np.ascontiguousarray(np.array(list(lscales)[:dim] + [1] * (3 - dim), np.int32))
        out = np.zeros(2 * int(np.prod(ii)))
        self.check(lib().dc_four_strucf_nufft(self.h, dim, len(xs[0]), _d(xyz[0]), _d(xyz[1]), _d(xyz[2]), float(occ),
                                              _i(ii), _i(ls), _d(out)))
        return out.view(np.complex128).reshape(tuple(int(v) for v in ii[:dim]), order="F")

    def four_run_nufft(self, desc, type_occ=None):
        """four_run_nufft: returns {'csf', 'dsi'} over inc (Fortran order)."""
        num = tuple(desc.inc)
        n = int(np.prod(num))
        csf = np.zeros(2 * n)
        dsi = np.zeros(n)
        occ = None if type_occ is None else np.ascontiguousarray(np.asarray(type_occ, np.float64).ravel())
        self.check(lib().dc_four_run_nufft(self.h, C.byref(desc), _d(occ), _d(csf), _d(dsi)))
        return {"csf": csf.view(np.complex128).reshape(num, order="F"), "dsi": dsi.reshape(num, order="F")}
