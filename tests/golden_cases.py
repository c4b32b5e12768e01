"""Loaders for the known-answer fixtures in tests/golden (written by
tests/golden/make_golden.py from closed forms and exact integer arithmetic)."""
import json
import os

import numpy as np

from diffusecode_amd import crystal as X

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CFPKT = 9999


def load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def structure_case(name):
    """(crystal, eck, vi, inc, cfact, F) of a structure-factor fixture; F is
    the expected complex F on the grid (Fortran order, abscissa fastest)."""
    d = load(name)
    sites = np.array(d["sites"], np.float64)
    types = np.array(d["types"], np.int32)
    pos, iscat = X.supercell(sites, types, tuple(d["cells"]))
    nscat = len(d["form_factor"])
    c = X.Crystal(pos=pos, iscat=iscat, icc=tuple(d["cells"]), ncatoms=len(sites), lattice=tuple(d["lattice"]),
                  scat=np.zeros((nscat, 11)), occ=np.ones(nscat), dw=np.zeros(nscat), name=name)
    cfact = np.ones((CFPKT + 1, nscat), np.complex128) * np.array(d["form_factor"], np.complex128)[None, :]
    inc = tuple(d["inc"])
    F = (np.array(d["F_re"]) + 1j * np.array(d["F_im"])).reshape(inc, order="F")
    return c, np.array(d["eck"], np.float64), np.array(d["vi"], np.float64), inc, cfact, F
