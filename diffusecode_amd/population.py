"""Population scoring on the MI355X library (dc_batch_chi2): a DIFFEV-style
generation of trial structures, each with its own atoms, form factors and
metric, scored by chi^2 against its own target grid in one pass.

DIFFEV runs one DISCUS macro per member over an MPI farm
(diffev/prog/diffev_kdo.f90:676, lib_f90/with_mpi.f90:281-609); here the
whole generation is one persistent GEMM launch plus one fused reduction and
one adjoint launch, replayed as a hipGraph from the second generation on.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .refine import scattering_table


class Population:
    """Members share the grid (eck, vi, inc in each member's reciprocal units)
    and the number of scattering types nscat.

    members: list of dicts with keys pos (N,3), types (N,), occ (N,),
    bdw (N,) or None, cfact (ncfact, nscat) complex, rten (3,3).
    """

    def __init__(self, members, eck, vi, inc, ctx=None, precision="bf16x3"):
        self.members = members
        self.eck = np.asarray(eck, float)
        self.vi = np.asarray(vi, float)
        self.inc = [int(v) for v in inc]
        self.ctx = ctx if ctx is not None else _lib.Context(precision=precision)
        sizes = [len(m["types"]) for m in members]
        self.atom_beg = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        self.types = np.concatenate([np.asarray(m["types"], np.int32) for m in members])
        self.cfact = np.stack([np.asarray(m["cfact"], np.complex128) for m in members], axis=2)
        self.rten = np.stack([np.asarray(m["rten"], float) for m in members])

    @classmethod
    def from_crystals(cls, crystals, eck, vi, inc, seed=0, **kw):
        """Members from crystals (crystal.population): per-atom occupancy and B
        from the type values with a seeded spread."""
        from .crystal import refinement_parameters
        members = []
        for k, cr in enumerate(crystals):
            occ, bdw = refinement_parameters(cr, seed=seed + k)
            members.append({"pos": cr.pos, "types": cr.iscat, "occ": occ, "bdw": bdw,
                            "cfact": scattering_table(cr.scat), "rten": cr.rten()})
        return cls(members, eck, vi, inc, **kw)

    @property
    def natoms(self):
        return int(self.atom_beg[-1])

    @property
    def npixels(self):
        return int(np.prod(self.inc))

    def params(self):
        pos = np.concatenate([m["pos"] for m in self.members])
        occ = np.concatenate([m["occ"] for m in self.members])
        bdw = None
        if any(m.get("bdw") is not None for m in self.members):
            bdw = np.concatenate([m["bdw"] if m.get("bdw") is not None else np.zeros(len(m["types"]))
                                  for m in self.members])
        return pos, occ, bdw

    def desc(self, pos=None, occ=None, bdw=None):
        p0, o0, b0 = self.params()
        pos = p0 if pos is None else pos
        occ = o0 if occ is None else occ
        bdw = b0 if bdw is None else bdw
        return _lib.Context.batch_desc(self.atom_beg, pos, self.types, occ, bdw, self.cfact, self.rten,
                                       self.eck, self.vi, self.inc)

    def score(self, target, weight=None, pos=None, occ=None, bdw=None, want_icalc=False, want_grad=True):
        """chi2 per member, I (P, M) if asked, grad (N, 5) if asked."""
        d, keep = self.desc(pos, occ, bdw)
        return self.ctx.batch_chi2(d, target, weight, want_icalc=want_icalc, want_grad=want_grad)
