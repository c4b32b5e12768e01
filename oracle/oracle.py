"""Python face of the CPU oracle (oracle/dc_oracle.c).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module as the checker.  The product package
(diffusecode_amd) never imports it.

Functions and what pins them (tests/test_oracle.py unless noted):

  strucf_exact   FP64 direct sum, h.r in long double reduced mod 1.
                 Pinned by: one atom at the origin gives F = 1 (x f(s));
                 perfect N1 x N2 x N3 lattice gives the Dirichlet product
                 prod_a sin(pi N_a h_a)/sin(pi h_a); rocksalt at integer hkl
                 gives 4(fA+fB)/4(fA-fB)/0; Friedel F(-h)=conj F(h);
                 |F|^2 invariant under integer translation; discrete
                 Parseval on a commensurate grid; brute force with Python
                 ``fractions`` + mpmath-free cmath on a few atoms.
  strucf_turbo   DISCUS's table algorithm (four_strucf_OMP_mod.f90:72-188).
                 Pinned by: its integer increments against exact rational
                 arithmetic (fractions) of NINT(2^22 frac(v.r)); its table
                 index against floor(iarg/64) mod 2^16 on hand-picked
                 values; closeness to strucf_exact within the table's
                 derived error bound (2^-16 truncation + 2^-23 per index
                 step); exact equality with strucf_exact at h=0; and bit for
                 bit against DISCUS's own four_strucf compiled from the
                 reference sources (ref_strucf below; tests/test_ref_strucf.py
                 and the vectors of tests/golden/ref_strucf_vectors.npz).
  ref_strucf / ref_fourier_lmn
                 Not restatements: DISCUS's four_strucf and fourier_lmn, built by
                 oracle/build_ref.py into oracle/_ref/ (present only where
                 the reference sources were available at build time).
  fourier_lmn    Reconstruction identity eck(:,1) = sum lmn(1:3) vi +
                 sum lmn(4:6) off(:,j) to 1e-7, and +-h phase identity; and
                 lmn / off equal to DISCUS's own fourier_lmn (ref_fourier_lmn,
                 tests/test_ref_strucf.py, tests/golden/ref_lmn_vectors.json).
  stltab         Closed form s = |h|/(2a) on a cubic metric, nint bin.
  formtab        Closed-form 4-Gaussian + exp(-B s^2) occupancy factor.
  test_friedel / apply_friedel
                 Hand-built grids (centred odd, off-centre, even), and the
                 full-grid exact sum equal to the mirrored half-grid sum.
  getatm / getav / csize / four_run
                 Perfect crystal with fave=100% and box lots gives dsi == 0
                 (Bragg subtraction KAT); LOT_OFF single lot gives
                 dsi = |F|^2; lot cell counts from celltoindex arithmetic;
                 cross-lot Welford variance against numpy two-pass variance.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
CFPKT = 9999
LOT_OFF, LOT_BOX, LOT_ELI = 1, 2, 3
MODE_EXACT, MODE_TURBO = 0, 1

_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
    return _lib


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _f64(a, n=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))
    if n is not None:
        assert a.size == n, (a.size, n)
    return a


def _i32(a, n=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.int32).ravel(order="F"))
    if n is not None:
        assert a.size == n
    return a


def nthreads_default():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def cexpt():
    out = np.empty(2 * 65536)
    lib().orc_cexpt(_d(out))
    return out.view(np.complex128)


def fourier_lmn(eck, vi, inc, rnd=(0, 0, 0, 0, 0, 0)):
    """eck (3,4), vi (3,3) Fortran-shaped arrays. Returns (status, lmn, off (3,3))."""
    eck = _f64(eck, 12)
    vi = _f64(vi, 9)
    inc = _i32(inc, 3)
    rnd = _f64(rnd, 6)
    lmn = np.zeros(6, np.int32)
    off = np.zeros(9)
    st = lib().orc_fourier_lmn(_d(eck), _d(vi), _i(inc), _d(rnd), _i(lmn), _d(off))
    return st, lmn, off.reshape(3, 3, order="F")


def stltab(xm, uin, vin, win, num, rten):
    num = _i32(num, 3)
    out = np.zeros(int(np.prod(num)), np.int32)
    st = lib().orc_stltab(_d(_f64(xm, 3)), _d(_f64(uin, 3)), _d(_f64(vin, 3)), _d(_f64(win, 3)),
                          _i(num), _d(_f64(rten, 9)), _i(out))
    return st, out.reshape(tuple(num), order="F")


def formtab(scat, occ, dw, ldbw=False, lxray=True, power=4, ano=False, delfr=None, delfi=None):
    scat = np.ascontiguousarray(np.asarray(scat, np.float64).reshape(-1, 11))
    nscat = scat.shape[0]
    occ = _f64(occ, nscat)
    dw = _f64(dw, nscat)
    delfr = _f64(np.zeros(nscat) if delfr is None else delfr, nscat)
    delfi = _f64(np.zeros(nscat) if delfi is None else delfi, nscat)
    out = np.zeros(2 * (CFPKT + 1) * nscat)
    lib().orc_formtab(C.c_int(nscat), _d(scat), C.c_int(int(lxray)), C.c_int(power), _d(occ), _d(dw),
                      C.c_int(int(ldbw)), C.c_int(int(ano)), _d(delfr), _d(delfi), _d(out))
    return out.view(np.complex128).reshape(CFPKT + 1, nscat, order="F")


def _xat(pos):
    pos = np.asarray(pos, np.float64)
    x = np.asfortranarray(pos.reshape(-1, 3))
    return x, x.shape[0]


def strucf_exact(pos, xm, uin, vin, win, num, fnum=None, cfact_col=None, istl=None, nthreads=None):
    """F over the grid xm + j uin + i vin + k win, j<fnum1, i<fnum2, k<fnum3.
    Returns complex128 (num1,num2,num3) Fortran-ordered array."""
    x, n = _xat(pos)
    num = _i32(num, 3)
    fnum = num if fnum is None else _i32(fnum, 3)
    lform = cfact_col is not None
    cf = _f64(np.zeros(2) if cfact_col is None else np.asarray(cfact_col, np.complex128).view(np.float64))
    ist = _i32(np.zeros(1) if istl is None else istl)
    out = np.zeros(2 * int(np.prod(num)))
    lib().orc_strucf_exact(_d(x), C.c_long(n), C.c_long(max(n, 1)), _d(_f64(xm, 3)), _d(_f64(uin, 3)),
                           _d(_f64(vin, 3)), _d(_f64(win, 3)), _i(num), _i(fnum), _d(cf), _i(ist),
                           C.c_int(int(lform)), _d(out), C.c_int(nthreads or nthreads_default()))
    return out.view(np.complex128).reshape(tuple(num), order="F")


def strucf_turbo(pos, uin, vin, win, off, lmn, num, fnum=None, cfact_col=None, istl=None, nthreads=None,
                 cex=None):
    x, n = _xat(pos)
    num = _i32(num, 3)
    fnum = num if fnum is None else _i32(fnum, 3)
    lform = cfact_col is not None
    cf = _f64(np.zeros(2) if cfact_col is None else np.asarray(cfact_col, np.complex128).view(np.float64))
    ist = _i32(np.zeros(1) if istl is None else istl)
    cex = cexpt() if cex is None else cex
    cexf = np.ascontiguousarray(np.asarray(cex, np.complex128)).view(np.float64)
    out = np.zeros(2 * int(np.prod(num)))
    lib().orc_strucf_turbo(_d(x), C.c_long(n), C.c_long(max(n, 1)), _d(_f64(uin, 3)), _d(_f64(vin, 3)),
                           _d(_f64(win, 3)), _d(_f64(off, 9)), _i(_i32(lmn, 6)), _i(num), _i(fnum),
                           _d(cexf), _d(cf), _i(ist), C.c_int(int(lform)), _d(out),
                           C.c_int(nthreads or nthreads_default()))
    return out.view(np.complex128).reshape(tuple(num), order="F")


_REF_PATH = os.path.join(_HERE, "_ref", "libref_strucf.so")
_ref = None


def ref_available() -> bool:
    """True when DISCUS's own kernel has been built into oracle/_ref
    (oracle/build_ref.py, from the reference sources)."""
    return os.path.exists(_REF_PATH)


def ref_strucf(pos, uin, vin, win, off, lmn, num, fnum=None, cfact_col=None, istl=None, nthreads=1, variant=0):
    """tcsf of DISCUS's own four_strucf (four_strucf_OMP_mod.f90:19-68),
    compiled from the reference sources (oracle/build_ref.py): one type,
    atoms pos (n, 3), the same arguments as strucf_turbo.  variant 0 = the
    dispatcher, 1 = four_strucf_serial, 2 = four_strucf_omp."""
    global _ref
    if _ref is None:
        # lazy binding: four GNU-extension and evaluator symbols of unreachable
        # file-information code stay unresolved (oracle/build_ref.py)
        _ref = C.CDLL(_REF_PATH, mode=os.RTLD_LAZY)
        _ref.ref_strucf.restype = None
    x = np.asfortranarray(np.asarray(pos, np.float64).reshape(-1, 3))
    n = x.shape[0]
    num = _i32(num, 3)
    fnum = num if fnum is None else _i32(fnum, 3)
    cf = np.ascontiguousarray(np.ones(1, np.complex128) if cfact_col is None else np.asarray(cfact_col, np.complex128))
    ist = np.asfortranarray(np.zeros(tuple(num), np.int32) if istl is None else
                            np.asarray(istl, np.int32).reshape(tuple(num), order="F"))
    out = np.zeros(tuple(num), np.complex128, order="F")
    _ref.ref_strucf(C.c_int(n), x.ctypes.data_as(_dp), _d(_f64(uin, 3)), _d(_f64(vin, 3)), _d(_f64(win, 3)),
                    _d(_f64(off, 9)), _i(_i32(lmn, 6)), _i(num), _i(fnum), C.c_int(cf.size),
                    cf.ctypes.data_as(_dp), ist.ctypes.data_as(C.POINTER(C.c_int)), C.c_int(int(cfact_col is not None)),
                    C.c_int(int(nthreads)), C.c_int(int(variant)), out.ctypes.data_as(_dp))
    return out


def ref_fourier_lmn(eck, vi, inc):
    """(ier_num, lmn, off) of DISCUS's own fourier_lmn
    (discus/prog/fourier_lmn_mod.f90:4-279), from the same build as
    ref_strucf."""
    global _ref
    if _ref is None:
        _ref = C.CDLL(_REF_PATH, mode=os.RTLD_LAZY)
        _ref.ref_strucf.restype = None
    _ref.ref_fourier_lmn.restype = C.c_int
    e = np.asfortranarray(np.asarray(eck, np.float64).reshape(3, 4))
    v = np.asfortranarray(np.asarray(vi, np.float64).reshape(3, 3))
    lmn = np.zeros(6, np.int32)
    off = np.zeros((3, 3), order="F")
    st = _ref.ref_fourier_lmn(e.ctypes.data_as(_dp), v.ctypes.data_as(_dp), _i(_i32(inc, 3)),
                              lmn.ctypes.data_as(C.POINTER(C.c_int)), off.ctypes.data_as(_dp))
    return int(st), lmn, off


def test_friedel(eck, vi, inc, ano=False, four_friedel=True):
    fnum = np.zeros(3, np.int32)
    ax = lib().orc_test_friedel(_d(_f64(eck, 12)), _d(_f64(vi, 9)), _i(_i32(inc, 3)), C.c_int(int(ano)),
                                C.c_int(int(four_friedel)), _i(fnum))
    return ax, fnum


def apply_friedel(num, fnum, acsf=None, csf=None, dsi=None):
    """In place on Fortran-ordered arrays (copies returned)."""
    num = _i32(num, 3)
    fnum = _i32(fnum, 3)
    outs = []
    ptrs = []
    for a, cplx in ((acsf, True), (csf, True), (dsi, False)):
        if a is None:
            outs.append(None)
            ptrs.append(None)
            continue
        b = np.asarray(a, np.complex128 if cplx else np.float64).ravel(order="F").copy()
        outs.append(b)
        ptrs.append((b.view(np.float64) if cplx else b).ctypes.data_as(_dp))
    lib().orc_apply_friedel(_i(num), _i(fnum), ptrs[0], ptrs[1], ptrs[2])
    return [None if b is None else b.reshape(tuple(num), order="F") for b in outs]


class _Run(C.Structure):
    _fields_ = [
        ("cr_pos", _dp), ("cr_iscat", _ip), ("natoms", C.c_long), ("icc", C.c_int * 3),
        ("ncatoms", C.c_int), ("dim0", C.c_double * 3), ("dim_lo", C.c_double * 3), ("nscat", C.c_int),
        ("eck", C.c_double * 12), ("vi", C.c_double * 9), ("inc", C.c_int * 3), ("rten", C.c_double * 9),
        ("cfact", _dp), ("ilots", C.c_int), ("ls", C.c_int * 3), ("nlots", C.c_int), ("lbeg", _ip),
        ("lperiod", C.c_int), ("fave", C.c_double), ("sel_cell", C.POINTER(C.c_ubyte)),
        ("four_friedel", C.c_int), ("ano", C.c_int), ("mode", C.c_int), ("rnd", C.c_double * 6),
        ("diff_scale", C.c_double),
    ]


def four_run(crystal, eck, vi, inc, cfact, ilots=LOT_OFF, ls=(1, 1, 1), lbeg=None, lperiod=True,
             fave=0.0, sel_cell=None, four_friedel=True, ano=False, mode=MODE_EXACT,
             rnd=(0, 0, 0, 0, 0, 0), diff_scale=1.0, nthreads=None, want_var=False):
    """Run-level restatement of four_run.  Returns dict csf, acsf, dsi
    (+ fmean, dvar when want_var) as Fortran-ordered (num1,num2,num3) arrays."""
    assert lib().orc_sizeof_run() == C.sizeof(_Run)
    pos = np.ascontiguousarray(np.asarray(crystal.pos, np.float64))
    isc = np.ascontiguousarray(np.asarray(crystal.iscat, np.int32))
    cf = np.asfortranarray(np.asarray(cfact, np.complex128))
    cff = cf.ravel(order="F").view(np.float64)
    nlots = 1 if ilots == LOT_OFF else len(lbeg)
    lb = np.ascontiguousarray(np.ones((nlots, 3), np.int32) if lbeg is None else np.asarray(lbeg, np.int32))
    sel = np.ascontiguousarray(np.zeros(1, np.uint8) if sel_cell is None else
                               np.asarray(sel_cell, np.uint8).ravel(order="F"))
    r = _Run()
    r.cr_pos = _d(pos)
    r.cr_iscat = _i(isc)
    r.natoms = crystal.natoms
    r.icc[:] = list(crystal.icc)
    r.ncatoms = crystal.ncatoms
    r.dim0[:] = list(crystal.dim0)
    r.dim_lo[:] = list(crystal.dim_lo)
    r.nscat = crystal.nscat
    r.eck[:] = list(_f64(eck, 12))
    r.vi[:] = list(_f64(vi, 9))
    r.inc[:] = [int(v) for v in inc]
    r.rten[:] = list(_f64(crystal.rten(), 9))
    r.cfact = cff.ctypes.data_as(_dp)
    r.ilots = ilots
    r.ls[:] = [int(v) for v in ls]
    r.nlots = nlots
    r.lbeg = _i(lb)
    r.lperiod = int(lperiod)
    r.fave = float(fave)
    r.sel_cell = sel.ctypes.data_as(C.POINTER(C.c_ubyte))
    r.four_friedel = int(four_friedel)
    r.ano = int(ano)
    r.mode = mode
    r.rnd[:] = list(_f64(rnd, 6))
    r.diff_scale = diff_scale
    num = tuple(int(v) for v in inc)
    ntot = int(np.prod(num))
    csf = np.zeros(2 * ntot)
    acsf = np.zeros(2 * ntot)
    dsi = np.zeros(ntot)
    fm = np.zeros(2 * ntot) if want_var else None
    dv = np.zeros(ntot) if want_var else None
    st = lib().orc_four_run(C.byref(r), _d(csf), _d(acsf), _d(dsi), None if fm is None else _d(fm),
                            None if dv is None else _d(dv), C.c_int(nthreads or nthreads_default()))
    out = {"status": st,
           "csf": csf.view(np.complex128).reshape(num, order="F"),
           "acsf": acsf.view(np.complex128).reshape(num, order="F"),
           "dsi": dsi.reshape(num, order="F")}
    if want_var:
        out["fmean"] = fm.view(np.complex128).reshape(num, order="F")
        out["dvar"] = dv.reshape(num, order="F")
    return out


# -- refinement model -----------------------------------------------------
def _model_args(pos, types, occ, bdw, cfact, eck, vi, inc, rten):
    p = np.ascontiguousarray(np.asarray(pos, np.float64).reshape(-1, 3))
    t = np.ascontiguousarray(np.asarray(types, np.int32).ravel())
    o = np.ascontiguousarray(np.asarray(occ, np.float64).ravel())
    b = None if bdw is None else np.ascontiguousarray(np.asarray(bdw, np.float64).ravel())
    cf = np.asfortranarray(np.asarray(cfact, np.complex128))
    cff = np.ascontiguousarray(cf.ravel(order="F")).view(np.float64)
    keep = (p, t, o, b, cff, _f64(eck, 12), _f64(vi, 9), _i32(inc, 3), _f64(rten, 9))
    args = [_d(p), _i(t), _d(o), None if b is None else _d(b), C.c_long(p.shape[0]), _d(cff), C.c_int(cf.shape[0]),
            _d(keep[5]), _d(keep[6]), _i(keep[7]), _d(keep[8])]
    return args, keep


def model_f(pos, types, occ, bdw, cfact, eck, vi, inc, rten, nthreads=None):
    """Exact FP64 F of the refinement model over the grid (Fortran order)."""
    args, keep = _model_args(pos, types, occ, bdw, cfact, eck, vi, inc, rten)
    num = tuple(int(v) for v in inc)
    out = np.zeros(2 * int(np.prod(num)))
    lib().orc_model_f(*args, _d(out), C.c_int(nthreads or nthreads_default()))
    return out.view(np.complex128).reshape(num, order="F")


def model_grad(pos, types, occ, bdw, cfact, eck, vi, inc, rten, F, G, nthreads=None):
    """Direct FP64 gradient (N, 5) of L with dL/dI = G, given the model's F."""
    args, keep = _model_args(pos, types, occ, bdw, cfact, eck, vi, inc, rten)
    f = np.ascontiguousarray(np.asarray(F, np.complex128).ravel(order="F")).view(np.float64)
    g = _f64(G)
    n = keep[0].shape[0]
    out = np.zeros(5 * n)
    lib().orc_model_grad(*args, _d(f), _d(g), _d(out), C.c_int(nthreads or nthreads_default()))
    return out.reshape(n, 5)


def model_fd(pos, types, occ, bdw, cfact, eck, vi, inc, rten, F0, target, weight, atoms, kinds, h=1e-6,
             nthreads=None):
    """FP64 central differences of chi^2 = sum w (|F|^2 - T)^2 for the listed
    (atom, kind) parameters, kind 0..4 = x, y, z, occ, B."""
    args, keep = _model_args(pos, types, occ, bdw, cfact, eck, vi, inc, rten)
    f = np.ascontiguousarray(np.asarray(F0, np.complex128).ravel(order="F")).view(np.float64)
    t = _f64(target)
    w = None if weight is None else _f64(weight)
    wp = None if w is None else _d(w)
    a = _i32(atoms)
    k = _i32(kinds, a.size)
    out = np.zeros(a.size)
    lib().orc_model_fd(*args, _d(f), _d(t), wp, C.c_int(a.size), _i(a), _i(k), C.c_double(h), _d(out),
                       C.c_int(nthreads or nthreads_default()))
    return out
