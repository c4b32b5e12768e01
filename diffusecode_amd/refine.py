"""Refinement model: intensity of a structure with per-atom parameters and its
reverse-mode derivative, on the MI355X library (dc_model_* of
include/diffuse_turbo.h), wrapped as ``torch.autograd.Function``.

REFINE (refine/prog/refine_run.f90:491-600) obtains the derivatives of the
computed intensity by re-running DISCUS once per parameter; here one adjoint
pass gives dL/d(x, y, z, occ, B) of every atom.  Parameters are the model's
own (per atom); the form-factor table carries f + f' + i f'' only
(four_formtab, fourier_sup.f90:2691-2743, with occupancy 1 and no B).
"""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import torch

from . import _lib
from .fourier import formtab


def scattering_table(scat, lxray=True, ano=False, delfr=None, delfi=None, power=4):
    """cfact(iq, t) = f_t(s) + f' + i f'' over the CFPKT+1 bins, (CFPKT+1, nscat)."""
    scat = np.asarray(scat, float).reshape(-1, 11)
    n = scat.shape[0]
    cr = SimpleNamespace(scat=scat, occ=np.ones(n), dw=np.zeros(n), nscat=n)
    return formtab(cr, lxray=lxray, ldbw=False, ano=ano, delfr=delfr, delfi=delfi, power=power)


class RefineModel:
    """Fixed part of a refinement problem: atom types, form factors, grid.

    types  (N,) 1-based scattering types (0 = vacancy)
    cfact  (ncfact, nscat) complex form factors without occupancy or B
    eck, vi, inc, rten  the grid and metric (grid_block / the fourier menu)
    """

    def __init__(self, types, cfact, eck, vi, inc, rten, precision="fp32", ctx=None):
        self.types = np.ascontiguousarray(np.asarray(types, np.int32).ravel())
        self.cfact = np.asarray(cfact, np.complex128)
        self.eck = np.asarray(eck, float)
        self.vi = np.asarray(vi, float)
        self.inc = [int(v) for v in inc]
        self.rten = np.asarray(rten, float)
        self.ctx = ctx if ctx is not None else _lib.Context(precision=precision)

    @property
    def shape(self):
        return tuple(self.inc)

    def desc(self, pos, occ, bdw):
        return _lib.Context.model_desc(pos, self.types, occ, bdw, self.cfact, self.eck, self.vi, self.inc,
                                       self.rten)

    # numpy face ---------------------------------------------------------------
    def intensity_np(self, pos, occ, bdw=None):
        d, keep = self.desc(pos, occ, bdw)
        return self.ctx.model_intensity(d)

    def vjp_np(self, pos, occ, bdw, fcalc, gi):
        d, keep = self.desc(pos, occ, bdw)
        return self.ctx.model_vjp(d, fcalc, gi)

    def jvp_np(self, pos, occ, bdw, fcalc, directions):
        """dI along directions (ndir, N, 5): the per-pixel derivative columns
        REFINE's least-squares loop needs (one per refined parameter)."""
        d, keep = self.desc(pos, occ, bdw)
        return self.ctx.model_jvp(d, fcalc, directions)

    def chi2_np(self, pos, occ, bdw, target, weight=None, want_grad=True):
        d, keep = self.desc(pos, occ, bdw)
        return self.ctx.model_chi2(d, target, weight, want_grad)

    # torch face ---------------------------------------------------------------
    def intensity(self, pos, occ, bdw):
        """I over the grid (float64 tensor of shape inc) with autograd."""
        return IntensityFn.apply(pos, occ, bdw, self)

    def chi2(self, pos, occ, bdw, target, weight=None):
        """chi^2 = sum w (I - target)^2 (0-d float64 tensor) with autograd."""
        return Chi2Fn.apply(pos, occ, bdw, self, target, weight)


def _np(t):
    return t.detach().to("cpu", torch.float64).numpy()


def _split_grad(g, like):
    gp = torch.from_numpy(np.ascontiguousarray(g[:, :3]))
    go = torch.from_numpy(np.ascontiguousarray(g[:, 3]))
    gb = torch.from_numpy(np.ascontiguousarray(g[:, 4]))
    return [x.to(device=t.device, dtype=t.dtype) for x, t in zip((gp, go, gb), like)]


class IntensityFn(torch.autograd.Function):
    """I(pos, occ, B); backward is the library's adjoint (dc_model_vjp)."""

    @staticmethod
    def forward(ctx, pos, occ, bdw, model):
        p, o, b = _np(pos), _np(occ), _np(bdw)
        icalc, fcalc = model.intensity_np(p, o, b)
        ctx.model = model
        ctx.params = (p, o, b)
        ctx.fcalc = fcalc
        ctx.like = (pos, occ, bdw)
        return torch.from_numpy(np.ascontiguousarray(icalc)).to(pos.device)

    @staticmethod
    def backward(ctx, gI):
        p, o, b = ctx.params
        gi = np.asfortranarray(_np(gI))
        g = ctx.model.vjp_np(p, o, b, ctx.fcalc, gi)
        gp, go, gb = _split_grad(g, ctx.like)
        return gp, go, gb, None


class Chi2Fn(torch.autograd.Function):
    """chi^2 and its gradient from one dc_model_chi2 call."""

    @staticmethod
    def forward(ctx, pos, occ, bdw, model, target, weight):
        t = np.asfortranarray(_np(target))
        w = None if weight is None else np.asfortranarray(_np(weight))
        chi, icalc, g = model.chi2_np(_np(pos), _np(occ), _np(bdw), t, w, want_grad=True)
        ctx.grads = _split_grad(g, (pos, occ, bdw))
        return torch.tensor(chi, dtype=torch.float64, device=pos.device)

    @staticmethod
    def backward(ctx, gout):
        gp, go, gb = ctx.grads
        s = gout.to(gp.dtype)
        return gp * s, go * s, gb * s, None, None, None
