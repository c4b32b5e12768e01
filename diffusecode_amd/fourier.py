"""Host mirror of DISCUS's ``fourier`` command set (discus/prog/fourier.f90).

The state and commands keep DISCUS's names and argument meaning so that a
``fourier`` macro translates line by line:

    DISCUS                         here
    ll -5,-5,0 / lr / ul / tl      f.ll(-5,-5,0) / f.lr(...) / f.ul(...) / f.tl(...)
    layer h1,..,h9, na, no         f.layer(ll, lr, ul, na, no)
    nabs 101 / nord / ntop         f.nabs(101) / f.nord(..) / f.ntop(..)
    lots box,8,8,8,64,y            f.lots("box", 8, 8, 8, 64, periodic=True)
    lots off                       f.lots("off")
    set aver,0.5                   f.set_aver(0.5)
    set symm,friedel|ignore        f.set_symm("friedel"|"ignore")
    temp use|ignore                f.temp("use"|"ignore")
    xray / neutron                 f.xray() / f.neutron()
    run                            f.run()
    calc h,k,l                     f.calc(h, k, l)

Results after ``run`` are ``csf``, ``acsf`` and ``dsi`` with DISCUS's
Fortran layout (num(1), num(2), num(3)), abscissa fastest
(discus/prog/diffuse_mod.f90:55-62).  Errors raise ``FourierError`` carrying
DISCUS's ier_num / ier_typ.

Random choices that DISCUS draws with ran1 (lot origins four_ranloc
fourier_sup.f90:2513-2530, the 'set aver' cell selection four_aver
:1233-1241) are drawn here from a seeded numpy generator, or passed in
explicitly, so the compute library itself never sees an RNG.
"""
from __future__ import annotations

import math

import numpy as np

from . import _lib

ER_COMM, ER_APPL, ER_FOUR = 1, 6, 9
CFPKT = 9999
CFINC = 0.001


class FourierError(RuntimeError):
    def __init__(self, ier_num, ier_typ, msg=""):
        super().__init__(f"ier_num={ier_num} ier_typ={ier_typ} {msg}")
        self.ier_num = ier_num
        self.ier_typ = ier_typ


def formtab(crystal, lxray=True, ldbw=False, ano=False, delfr=None, delfi=None, power=4):
    """four_formtab (fourier_sup.f90:2691-2743) with form (:2947-2971):
    cfact(iq, t) = sb (f + f') + i sb f'',  sb = occ exp(-B s^2) if ldbw else occ,
    s = iq * CFINC.  Returns complex128 (CFPKT+1, nscat), Fortran order."""
    nscat = crystal.nscat
    q2 = (np.arange(CFPKT + 1, dtype=np.float64) * CFINC) ** 2
    out = np.zeros((CFPKT + 1, nscat), dtype=np.complex128, order="F")
    for t in range(nscat):
        sc = crystal.scat[t]
        f = np.full_like(q2, sc[0])
        if lxray:
            for i in range(1, power + 1):
                f = f + sc[2 * i - 1] * np.exp(-sc[2 * i] * q2)
        fp = delfr[t] if (ano and delfr is not None) else 0.0
        fpp = delfi[t] if (ano and delfi is not None) else 0.0
        sb = np.exp(-(crystal.dw[t] * q2)) * crystal.occ[t] if ldbw else np.full_like(q2, crystal.occ[t])
        out[:, t] = sb * (f + fp) + 1j * sb * fpp
    return out


class Fourier:
    """State of the fourier menu (defaults of diffuse_mod.f90 / fourier_reset)."""

    def __init__(self, crystal=None, precision="fp32", ctx=None, seed=0):
        self.eck = np.array([[0.0, 0.0, 0.0], [5.0, 0.0, 0.0], [0.0, 5.0, 0.0], [0.0, 0.0, 0.0]]).T.copy()
        self.inc = [121, 121, 1]
        self.divis = [float(max(1, self.inc[0] - 1)), float(max(1, self.inc[1] - 1)), 1.0]
        self.vi = np.zeros((3, 3))
        self._update_vi()
        self.ilots = 1
        self.nlots = 1
        self.ls_xyz = [5, 5, 5]
        self.lperiod = True
        self.lot_all = False
        self.fave = 0.0
        self.four_friedel = True
        self.ldbw = False
        self.lxray = True
        self.ano = False
        self.diff_scale = 1.0
        self.ltop_c = False
        self.ltop_n = False
        self.rng = np.random.default_rng(seed)
        # 'set technique/filter/scale/damp/width/symm/mode' (fourier.f90:1090-1200, fourier_reset.f90:100-105)
        self.four_tech = "turbo"
        self.four_filter = False
        self.four_rscale = 1.0
        self.four_damp = 0.5
        self.four_width = 4
        self.four_symm = False
        self.symm_ops = None            # point-group operations (identity first) for 'symm apply'
        self.four_accum = "single"
        self.res_sigma = [0.0, 0.0, 0.0]   # diff_res(1, 1:3)
        self.res_tr = np.zeros((3, 3))     # diff_tr
        self.type_occ = None               # cr_occ per type, the NUFFT weight
        self.want_dsi3d = False            # four_weight's dsi3d (for 3D-PDF output)
        self.extr = [1, 2, 3]              # extr_abs/_ord/_top: h, k, l along the axes (graphic.f90:2695-2697)
        self.ctx = ctx
        self.precision = precision
        self.crystal = None
        self.csf = self.acsf = self.dsi = None
        if crystal is not None:
            self.set_crystal(crystal)

    # -- crystal / context --------------------------------------------------
    def set_crystal(self, crystal):
        if self.ctx is None:
            self.ctx = _lib.Context(precision=self.precision)
        self.ctx.crystal_set(crystal)
        self.crystal = crystal

    # -- corners and increments (fourier.f90:402-466, 532-621) --------------
    def _update_vi(self):
        for i in range(3):
            self.vi[i, 0] = (self.eck[i, 1] - self.eck[i, 0]) / self.divis[0]
            self.vi[i, 1] = (self.eck[i, 2] - self.eck[i, 0]) / self.divis[1]
            self.vi[i, 2] = (self.eck[i, 3] - self.eck[i, 0]) / self.divis[2]

    def _corner(self, j, h, k, l):
        self.eck[:, j] = (h, k, l)
        self._update_vi()

    def ll(self, h, k, l):
        self._corner(0, h, k, l)

    def lr(self, h, k, l):
        self._corner(1, h, k, l)

    def ul(self, h, k, l):
        self._corner(2, h, k, l)

    def tl(self, h, k, l):
        self.ltop_c = True
        self._corner(3, h, k, l)

    def layer(self, ll, lr, ul, na, no):
        for j, v in enumerate((ll, lr, ul)):
            self.eck[:, j] = v
        self.eck[:, 3] = self.eck[:, 0]
        self.inc = [int(round(na)), int(round(no)), 1]
        self.divis = [float(max(1, self.inc[0] - 1)), float(max(1, self.inc[1] - 1)), 1.0]
        self._update_vi()

    def _n(self, axis, n):
        if n <= 0:
            raise FourierError(-12, ER_APPL, "number of points must be positive")
        self.inc[axis] = int(round(n))
        self.divis[axis] = float(max(1, self.inc[axis] - 1))
        self._update_vi()

    def nabs(self, n):
        self._n(0, n)

    def nord(self, n):
        self._n(1, n)

    def ntop(self, n):
        self._n(2, n)
        self.ltop_n = True

    # -- lots, aver, symmetry, radiation (fourier.f90:467-531, 1066-1247) ----
    def lots(self, shape, nx=5, ny=5, nz=5, nlots=1, periodic=True):
        shape = shape.lower()
        if shape.startswith("o"):
            self.ilots, self.nlots = 1, 1
            return
        if shape[0] not in "be":
            raise FourierError(-2, ER_FOUR, "unknown lot shape")
        icc = self.crystal.icc if self.crystal is not None else (10 ** 9,) * 3
        for v, m in zip((nx, ny, nz), icc):
            if not (0 < v <= m):
                raise FourierError(-101, ER_APPL, "lot size outside crystal")
        self.ilots = 2 if shape[0] == "b" else 3
        self.ls_xyz = [int(nx), int(ny), int(nz)]
        self.lperiod = bool(periodic)
        self.lot_all = isinstance(nlots, str) and nlots.lower().startswith("all")
        if self.lot_all:
            self.nlots = int(np.prod(self.csize()))
        else:
            self.nlots = int(nlots)

    def csize(self):
        """four_csize (fourier_sup.f90:2533-2555)."""
        icc = self.crystal.icc
        if self.lperiod:
            return list(icc)
        return [max(1, icc[i] - self.ls_xyz[i] + 1) for i in range(3)]

    def set_aver(self, fraction):
        self.fave = float(fraction)

    def set_symm(self, mode):
        self.four_friedel = mode.lower().startswith("f")

    def temp(self, mode):
        self.ldbw = mode.lower().startswith("u")

    def xray(self):
        self.lxray = True

    def neutron(self):
        self.lxray = False

    def set_options(self, technique=None, filter=None, scale=None, damp=None, width=None, symm=None, mode=None):
        """Optional parameters of 'set' (fourier.f90:1090-1200): technique
        turbo|nufft, filter lanczos|off, scale (four_rscale with a filter,
        else diff_scale), damp, width, symm apply|friedel|ignore, mode
        single|init|add|finished."""
        if symm is not None:
            m = symm.lower()
            if m not in ("apply", "friedel", "ignore"):
                raise FourierError(-6, ER_COMM, "Symmetry paramter must be 'apply', 'friedel' or 'ignore'")
            self.four_symm = m == "apply"
            self.four_friedel = m != "ignore"
        if technique is not None:
            if technique.lower() not in ("turbo", "nufft"):
                raise FourierError(-6, ER_COMM, "Fourier technique must be 'turbo' or 'nufft'")
            self.four_tech = technique.lower()
        if filter is not None:
            self.four_filter = filter.lower() == "lanczos"
        if scale is not None:
            if filter is not None and filter.lower() == "lanczos":
                self.four_rscale = float(scale)
            elif filter is None:
                self.diff_scale = float(scale)
        if damp is not None:
            self.four_damp = float(damp)
        if width is not None:
            self.four_width = int(round(width))
        if mode is not None:
            m = mode.lower()
            if m not in ("single", "init", "add", "finished"):
                raise FourierError(-6, ER_COMM, "mode must be single, init, add or finished")
            self.four_accum = m

    def resolution(self, sigabs=0.0, sigord=0.0, sigtop=0.0, tr=None):
        """Resolution of four_conv ('run' options sigabs/sigord/sigtop, four_resolution,
        fourier_conv.f90:426-650) with the default resolution vectors (the
        increment vectors): simat's columns are then vi(:,k) sigma_k / |vi(:,k)|,
        so diff_tr = simat^-1 vi = diag(|vi(:,k)| / sigma_k) on the non-flat axes.
        tr: a full diff_tr (3x3) for other resolution vectors."""
        self.res_sigma = [float(sigabs), float(sigord), float(sigtop)]
        if tr is not None:
            self.res_tr = np.asarray(tr, float).reshape(3, 3)
            return
        g = self.crystal.rten() if self.crystal is not None else np.eye(3)
        self.res_tr = np.zeros((3, 3))
        for k in range(3):
            if self.inc[k] > 1 and self.res_sigma[k] > 0.0:
                v = self.vi[:, k]
                self.res_tr[k, k] = np.sqrt(v @ g @ v) / self.res_sigma[k]

    def _extr(self, axis, comp):
        """'abscissa | ordinate | top  h|k|l' (fourier.f90:223-240, 622-639,
        889-906): which reciprocal component runs along the axis."""
        m = {"h": 1, "k": 2, "l": 3}
        if comp not in m:
            raise FourierError(-6, ER_COMM, "parameter must be h, k or l")
        self.extr[axis] = m[comp]

    def abscissa(self, comp):
        self._extr(0, comp)

    def ordinate(self, comp):
        self._extr(1, comp)

    def top(self, comp):
        self._extr(2, comp)

    def output_hdf5(self, path, value="inten", laver=False, scale=1.0, valmax=0.0, **qval_kw):
        """DISCUS output with 'outfile type hdf5' (graphic.f90:606-636): the
        qval of every pixel of the last run (dc_qval on the GPU) times the output
        scale, written as a Yell 1.0 file (dc_write_yell, hdf_write.f90:12-326).
        Returns the values written (before valmax)."""
        from . import yell
        if self.dsi is None or self.crystal is None:
            raise FourierError(-118, ER_APPL, "You need to calculate a Fourier / Patterson first")
        c = self.crystal
        qd, keep = _lib.Context.qval_desc(self.inc, self.eck, self.vi, c.rten(), value, laver=laver, **qval_kw)
        try:
            q = self.ctx.qval(qd, csf=self.csf, acsf=self.acsf, dsi=self.dsi)
        except _lib.DCError as e:
            raise FourierError(e.ier_num, e.ier_typ, str(e)) from None
        q = q * scale
        yell.write_yell(path, q, self.eck, self.vi, c.lattice[:3], c.lattice[3:], value=qd.value, valmax=valmax,
                        extr=self.extr)
        return q

    def set_precision(self, precision):
        self.precision = precision
        if self.ctx is not None:
            self.ctx.set_precision(precision)

    # -- run (fourier.f90:639-747) -------------------------------------------
    def lot_origins(self):
        """Lot origins: enumerated for 'all' (fourier_sup.f90:144-150), else random cells."""
        cs = self.csize()
        if self.lot_all:
            out = []
            for n in range(self.nlots):
                ii = n
                l3 = ii // (cs[0] * cs[1]) + 1
                ii -= (l3 - 1) * cs[0] * cs[1]
                l2 = ii // cs[0] + 1
                ii -= (l2 - 1) * cs[0]
                out.append((ii % cs[0] + 1, l2, l3))
            return np.array(out, np.int32)
        return (np.floor(self.rng.random((self.nlots, 3)) * np.array(cs)).astype(np.int32) + 1)

    def aver_cells(self):
        """'set aver' cell selection: each cell with probability fave (four_aver :1235-1241)."""
        cs = self.csize()
        return (self.rng.random((cs[0], cs[1], cs[2])) <= self.fave).astype(np.uint8)

    def _prepare(self):
        c = self.crystal
        if c is None:
            raise FourierError(-198, ER_APPL, "No fourier will be calculated")
        if c.natoms < 1:
            raise FourierError(-198, ER_APPL, "No fourier will be calculated")
        if not self.ltop_c:
            if (not self.ltop_n) or self.inc[2] == 1:
                self.eck[:, 3] = self.eck[:, 0]
                self.vi[:, 2] = 0.0
                self.inc[2] = 1
                self.divis[2] = 1.0
            else:
                raise FourierError(-191, ER_APPL, "three-D corner not defined")

    def run(self, lbeg=None, sel_cell=None, want_var=False, device_out=None):
        self._prepare()
        c = self.crystal
        cfact = formtab(c, lxray=self.lxray, ldbw=self.ldbw, ano=self.ano)
        if self.four_tech == "nufft" and device_out is None:
            return self._run_nufft(cfact)
        if self.ilots != 1 and lbeg is None:
            lbeg = self.lot_origins()
        if self.fave != 0.0 and sel_cell is None:
            sel_cell = self.aver_cells()
        desc, keep = _lib.Context.run_desc(self.eck, self.vi, self.inc, c.rten(), cfact, ilots=self.ilots,
                                           ls=self.ls_xyz, lbeg=lbeg, lperiod=self.lperiod, fave=self.fave,
                                           sel_cell=sel_cell, four_friedel=self.four_friedel, ano=self.ano,
                                           diff_scale=self.diff_scale)
        self.last_lbeg, self.last_sel_cell = lbeg, sel_cell
        try:
            if device_out is not None:
                st = self.ctx.four_run_dev(desc, **device_out)
                self.status = st
                return device_out
            out = self.ctx.four_run(desc, want_var=want_var)
        except _lib.DCError as e:
            raise FourierError(e.ier_num, e.ier_typ, str(e)) from None
        self.status = out["status"]
        self.csf, self.acsf, self.dsi = out["csf"], out["acsf"], out["dsi"]
        if want_var:
            self.fmean, self.dvar = out["fmean"], out["dvar"]
        self._tail(out)
        return out

    def _run_nufft(self, cfact):
        """four_run_nufft (fourier_sup.f90:327-690) up to dsi, then the tail."""
        c = self.crystal
        desc, keep = _lib.Context.run_desc(self.eck, self.vi, self.inc, c.rten(), cfact, fave=self.fave,
                                           diff_scale=self.diff_scale)
        try:
            out = self.ctx.four_run_nufft(desc, self.type_occ)
        except _lib.DCError as e:
            raise FourierError(e.ier_num, e.ier_typ, str(e)) from None
        out["status"] = 0
        out["acsf"] = np.zeros_like(out["csf"])
        self.status = 0
        self.csf, self.acsf, self.dsi = out["csf"], out["acsf"], out["dsi"]
        self._tail(out)
        return out

    def _tail(self, out):
        """four_weight, filter, four_conv, filter of dsi3d (fourier_sup.f90:279-290)
        and, in 'single' mode with 'symm apply', the point-group average
        (four_accumulate :3784-3840)."""
        c = self.crystal
        conv = any(sv >= 1.0e-6 for sv in self.res_sigma)
        self.dsi3d = None
        if self.four_filter or conv or self.want_dsi3d:
            vr = float(np.sqrt(np.linalg.det(c.rten())))
            pd = _lib.Context.post_desc(self.inc, self.eck, self.vi, c.rten(), vr, fave=self.fave, nlots=self.nlots,
                                        icc=c.icc, ls=self.ls_xyz, lots_off=self.ilots == 1,
                                        diff_scale=self.diff_scale, res_sigma=self.res_sigma,
                                        res_tr=self.res_tr.ravel(order="F"), lanczos=self.four_filter,
                                        rscale=self.four_rscale, damp=self.four_damp, width=self.four_width)
            try:
                csf, dsi, d3 = self.ctx.four_post(pd, out["csf"], out["dsi"], want_dsi3d=self.want_dsi3d)
            except _lib.DCError as e:
                raise FourierError(e.ier_num, e.ier_typ, str(e)) from None
            out["csf"], out["dsi"], out["dsi3d"] = csf, dsi, d3
            self.csf, self.dsi, self.dsi3d = csf, dsi, d3
        if self.four_accum == "single" and self.four_symm and self.symm_ops is not None:
            sd, keep = _lib.Context.symm_desc(self.inc, self.eck, self.vi, fave=self.fave, four_symm=True,
                                              ops=self.symm_ops)
            csf, dsi = self.ctx.four_fill(sd, out["csf"], out["dsi"])
            out["csf"], out["dsi"] = csf, dsi
            self.csf, self.dsi = csf, dsi

    def calc(self, h, k, l):
        """Single reflection (fourier.f90:243-294 -> calc_000): a 1-point grid."""
        saved = (self.eck.copy(), self.vi.copy(), list(self.inc), list(self.divis), self.ltop_c, self.ilots, self.nlots)
        try:
            self.eck[:, :] = np.array([h, k, l], float)[:, None]
            self.inc = [1, 1, 1]
            self.divis = [1.0, 1.0, 1.0]
            self.vi[:] = 0.0
            self.ltop_c = True
            self.ilots, self.nlots = 1, 1
            out = self.run()
            return complex(out["csf"].ravel()[0]), float(out["dsi"].ravel()[0])
        finally:
            self.eck, self.vi, self.inc, self.divis, self.ltop_c, self.ilots, self.nlots = saved


def pairs_full_grid(crystal, inc, ilots=1, ls=None, nlots=1):
    """Work unit: atom.hkl pairs = sum_lots sum_types N_type,lot x P_full."""
    P = int(np.prod(inc))
    if ilots == 1:
        n = int(np.count_nonzero(crystal.iscat > 0))
        return n * P
    per_lot = int(np.prod(ls)) * crystal.ncatoms
    frac = np.count_nonzero(crystal.iscat > 0) / max(1, crystal.natoms)
    return int(round(per_lot * frac)) * P * nlots


def grid_block(lo, hi, n):
    """eck/vi/inc for a block lo..hi (each a 3-vector) with n points per axis;
    axes with n=1 are flat."""
    lo = np.asarray(lo, float)
    hi = np.asarray(hi, float)
    eck = np.zeros((3, 4))
    eck[:, 0] = lo
    for a in range(3):
        v = lo.copy()
        v[a] = hi[a]
        eck[:, a + 1] = v
    inc = [int(x) for x in n]
    vi = np.zeros((3, 3))
    for a in range(3):
        vi[:, a] = (eck[:, a + 1] - eck[:, 0]) / max(1, inc[a] - 1)
    return eck, vi, inc
