"""Seeded synthetic crystals in DISCUS's data model.

A crystal is held the way DISCUS's ``crystal_mod`` holds it
(discus/prog/crystal_mod.f90:25-138): fractional positions ``cr_pos(3,N)``,
1-based scattering types ``cr_iscat`` (0 marks a vacancy/VOID, never
gathered because ``four_getatm`` compares against types >= 1,
discus/prog/fourier_sup.f90:1842,1878), the supercell size ``cr_icc``, sites
per cell ``cr_ncatoms`` and the cell-major atom order of ``celltoindex``
(discus/prog/celltoindex.f90:21):

    iatom = ((k*icc1*icc2 + j*icc1 + i) * ncatoms + site)      (0-based)

Scattering is the X-ray IAM model of ``four_formtab``/``form``
(fourier_sup.f90:2691-2743, 2947-2971): per type 11 coefficients
``c, a1, b1, ..., a4, b4`` (plus padding), occupancy and isotropic B.

Everything is seeded and built with numpy; nothing here touches a GPU.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np


@dataclasses.dataclass
class Crystal:
    pos: np.ndarray            # (N, 3) float64 fractional coordinates
    iscat: np.ndarray          # (N,) int32, 1-based types, 0 = vacancy
    icc: tuple                 # supercell size in unit cells
    ncatoms: int               # sites per unit cell
    lattice: tuple             # a, b, c (A), alpha, beta, gamma (deg)
    scat: np.ndarray           # (nscat, 11) c, a1, b1, ..., a4, b4, 0, 0
    occ: np.ndarray            # (nscat,) per-type occupancy
    dw: np.ndarray             # (nscat,) per-type isotropic B (A^2)
    dim0: tuple = (0.0, 0.0, 0.0)   # cr_dim0(:,1): origin of cell (1,1,1)
    name: str = "crystal"

    @property
    def natoms(self) -> int:
        return int(self.pos.shape[0])

    @property
    def nscat(self) -> int:
        return int(self.scat.shape[0])

    @property
    def dim_lo(self):
        """cr_dim(:,1), the lowest fractional coordinate (cell-origin frame)."""
        return tuple(float(v) for v in self.dim0)

    def rten(self) -> np.ndarray:
        """Reciprocal metric tensor G* (cr_rten) as a (3,3) array."""
        return reciprocal_metric(*self.lattice)

    def type_counts(self) -> np.ndarray:
        return np.bincount(self.iscat, minlength=self.nscat + 1)[1:]


def reciprocal_metric(a, b, c, alpha=90.0, beta=90.0, gamma=90.0) -> np.ndarray:
    ca, cb, cg = (math.cos(math.radians(x)) for x in (alpha, beta, gamma))
    g = np.array([[a * a, a * b * cg, a * c * cb],
                  [a * b * cg, b * b, b * c * ca],
                  [a * c * cb, b * c * ca, c * c]], dtype=np.float64)
    return np.linalg.inv(g)


def random_form_factors(rng: np.random.Generator, nscat: int) -> np.ndarray:
    """f(s) = sum_{i=1..4} a_i exp(-b_i s^2) + c, a_i~U[1,5], b_i~U[0.5,30], c~U[0,2]."""
    scat = np.zeros((nscat, 11), dtype=np.float64)
    for t in range(nscat):
        scat[t, 0] = rng.uniform(0.0, 2.0)
        for i in range(4):
            scat[t, 1 + 2 * i] = rng.uniform(1.0, 5.0)
            scat[t, 2 + 2 * i] = rng.uniform(0.5, 30.0)
    return scat


def _cell_grid(icc):
    n1, n2, n3 = icc
    k, j, i = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
    # celltoindex order: i fastest, then j, then k
    return np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64)


def supercell(sites: np.ndarray, site_types: np.ndarray, icc) -> tuple:
    """Positions and types of a supercell in DISCUS atom order."""
    cells = _cell_grid(icc)
    ns = sites.shape[0]
    pos = (cells[:, None, :] + sites[None, :, :]).reshape(-1, 3)
    types = np.tile(site_types, cells.shape[0]).astype(np.int32)
    assert pos.shape[0] == cells.shape[0] * ns
    return pos, types


ROCKSALT_SITES = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                           [.5, 0, 0], [0, .5, 0], [0, 0, .5], [.5, .5, .5]], dtype=np.float64)
ROCKSALT_TYPES = np.array([1, 1, 1, 1, 2, 2, 2, 2], dtype=np.int32)
FCC_SITES = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]], dtype=np.float64)
BCC_SITES = np.array([[0, 0, 0], [.5, .5, .5]], dtype=np.float64)


def rocksalt(n=10, a=5.64, vacancy=0.05, sigma=0.05, seed=1) -> Crystal:
    """NaCl-type n^3 supercell, random vacancies and Gaussian displacements (A)."""
    rng = np.random.default_rng(seed)
    pos, types = supercell(ROCKSALT_SITES, ROCKSALT_TYPES, (n, n, n))
    vac = rng.random(types.shape[0]) < vacancy
    types[vac] = 0
    pos = pos + rng.normal(0.0, sigma / a, size=pos.shape)
    scat = random_form_factors(rng, 2)
    return Crystal(pos, types, (n, n, n), 8, (a, a, a, 90.0, 90.0, 90.0), scat,
                   np.ones(2), np.zeros(2), name=f"rocksalt{n}")


def _fcc_neighbours(icc):
    """Nearest-neighbour table (12 per site) of an fcc supercell, periodic."""
    n1, n2, n3 = icc
    nsite = 4
    offs = []
    for s in range(nsite):
        for t in range(nsite):
            for d in np.ndindex(3, 3, 3):
                dv = np.array(d) - 1 + FCC_SITES[t] - FCC_SITES[s]
                if abs(np.dot(dv, dv) - 0.5) < 1e-9:
                    offs.append((s, t, d[0] - 1, d[1] - 1, d[2] - 1))
    cells = _cell_grid(icc).astype(np.int64)
    ncell = cells.shape[0]
    nb = np.zeros((ncell * nsite, 12), dtype=np.int64)
    fill = np.zeros(ncell * nsite, dtype=np.int64)
    for (s, t, dx, dy, dz) in offs:
        ci = (cells[:, 0] + dx) % n1
        cj = (cells[:, 1] + dy) % n2
        ck = (cells[:, 2] + dz) % n3
        tgt = ((ck * n2 + cj) * n1 + ci) * nsite + t
        src = np.arange(ncell) * nsite + s
        nb[src, fill[src]] = tgt
        fill[src] += 1
    assert np.all(fill == 12)
    return nb


def ising_swap_mc(types: np.ndarray, nb: np.ndarray, steps: int, beta_j: float, seed: int) -> np.ndarray:
    """Seeded Metropolis swap Monte Carlo with energy J * sum_nn delta(t_i, t_j).

    J > 0 (beta_j > 0) penalises like neighbours, which builds Warren-Cowley
    short-range order of the alternating kind.  Composition is conserved.
    """
    rng = np.random.default_rng(seed)
    t = types.copy()
    n = t.shape[0]
    picks = rng.integers(0, n, size=(steps, 2))
    draws = rng.random(steps)
    for s in range(steps):
        a, b = picks[s]
        ta, tb = t[a], t[b]
        if ta == tb:
            continue
        na, nbb = nb[a], nb[b]
        # like-pair counts before and after swapping a<->b
        before = np.count_nonzero(t[na] == ta) + np.count_nonzero(t[nbb] == tb)
        t[a], t[b] = tb, ta
        after = np.count_nonzero(t[na] == tb) + np.count_nonzero(t[nbb] == ta)
        d = after - before
        if d > 0 and draws[s] >= math.exp(-beta_j * d):
            t[a], t[b] = ta, tb
    return t


def fcc_ordered(n=40, a=4.0, ntypes=3, mc_steps=100_000, beta_j=0.7, seed=2) -> Crystal:
    """fcc n^3 supercell with ntypes species and swap-MC short-range order."""
    rng = np.random.default_rng(seed)
    pos, _ = supercell(FCC_SITES, np.zeros(4, np.int32), (n, n, n))
    nat = pos.shape[0]
    types = (np.arange(nat) % ntypes + 1).astype(np.int32)
    rng.shuffle(types)
    if mc_steps > 0:
        types = ising_swap_mc(types, _fcc_neighbours((n, n, n)), mc_steps, beta_j, seed + 1000)
    scat = random_form_factors(rng, ntypes)
    return Crystal(pos, types.astype(np.int32), (n, n, n), 4, (a, a, a, 90.0, 90.0, 90.0), scat,
                   np.ones(ntypes), np.zeros(ntypes), name=f"fcc{n}")


def _smooth_noise(rng, shape, width):
    """Gaussian-filtered white noise on a periodic grid, unit variance."""
    w = rng.normal(size=shape)
    k = [np.fft.fftfreq(s) for s in shape]
    kk = sum(np.meshgrid(*[x * x for x in k], indexing="ij"))
    g = np.exp(-2.0 * (math.pi * width) ** 2 * kk)
    out = np.real(np.fft.ifftn(np.fft.fftn(w) * g))
    return out / out.std()


def fcc_correlated(n=100, a=4.0, subst=0.10, sigma=0.08, corr_cells=2.0, seed=3) -> Crystal:
    """fcc n^3, 2 types with random substitution and spatially correlated displacements."""
    rng = np.random.default_rng(seed)
    pos, types = supercell(FCC_SITES, np.ones(4, np.int32), (n, n, n))
    types[rng.random(types.shape[0]) < subst] = 2
    disp = np.empty_like(pos)
    for c in range(3):
        field = _smooth_noise(rng, (n, n, n), corr_cells)   # indexed [k, j, i]
        disp[:, c] = np.repeat(field.ravel(), 4)
    pos = pos + disp * (sigma / a)
    scat = random_form_factors(rng, 2)
    return Crystal(pos, types, (n, n, n), 4, (a, a, a, 90.0, 90.0, 90.0), scat,
                   np.ones(2), np.zeros(2), name=f"fcc{n}c")


def orthorhombic(ncell=(16, 16, 16), abc=(4.0, 5.0, 6.0), nsite=8, ntypes=4, seed=4) -> Crystal:
    """Orthorhombic cell with nsite seeded random sites (32,768 atoms by default)."""
    rng = np.random.default_rng(seed)
    sites = rng.random((nsite, 3))
    site_types = (np.arange(nsite) % ntypes + 1).astype(np.int32)
    pos, types = supercell(sites, site_types, ncell)
    pos = pos + rng.normal(0.0, 0.01, size=pos.shape)
    scat = random_form_factors(rng, ntypes)
    occ = rng.uniform(0.7, 1.0, size=ntypes)
    dw = rng.uniform(0.3, 1.5, size=ntypes)
    return Crystal(pos, types, tuple(ncell), nsite, (abc[0], abc[1], abc[2], 90.0, 90.0, 90.0),
                   scat, occ, dw, name="ortho")


def refinement_parameters(cr: Crystal, seed=6, occ_spread=0.1, b_sigma=0.05):
    """Per-atom occupancies and isotropic B of a crystal (REFINE's parameter
    vector): the type values of cr.occ / cr.dw with seeded per-atom spread."""
    rng = np.random.default_rng(seed)
    t = np.maximum(cr.iscat, 1) - 1
    occ = cr.occ[t] * rng.uniform(1.0 - occ_spread, 1.0, size=cr.natoms)
    bdw = np.maximum(cr.dw[t] + rng.normal(0.0, b_sigma, size=cr.natoms), 0.05)
    return occ, bdw


def perturb_parameters(pos, occ, bdw, seed=7, dx=0.002, docc=0.02, db=0.05):
    """Seeded perturbation (delta r, delta occ, delta B) of a parameter set."""
    rng = np.random.default_rng(seed)
    return (pos + rng.normal(0.0, dx, size=pos.shape),
            occ + rng.normal(0.0, docc, size=occ.shape),
            np.maximum(bdw + rng.normal(0.0, db, size=bdw.shape), 0.01))


def population(nmembers=96, nmin=2000, nmax=60000, seed=5):
    """DIFFEV-like population: fcc/bcc/rocksalt supercells of log-uniform size."""
    rng = np.random.default_rng(seed)
    out = []
    kinds = [("fcc", FCC_SITES, np.array([1, 2, 1, 2], np.int32), 4.0),
             ("bcc", BCC_SITES, np.array([1, 2], np.int32), 3.2),
             ("rocksalt", ROCKSALT_SITES, ROCKSALT_TYPES, 5.64)]
    for m in range(nmembers):
        target = math.exp(rng.uniform(math.log(nmin), math.log(nmax)))
        name, sites, stypes, a = kinds[m % 3]
        n = max(2, int(round((target / sites.shape[0]) ** (1.0 / 3.0))))
        pos, types = supercell(sites, stypes, (n, n, n))
        types[rng.random(types.shape[0]) < 0.03] = 0
        pos = pos + rng.normal(0.0, 0.03 / a, size=pos.shape)
        out.append(Crystal(pos, types, (n, n, n), sites.shape[0], (a, a, a, 90.0, 90.0, 90.0),
                           random_form_factors(rng, 2), np.ones(2), np.zeros(2),
                           name=f"{name}{n}"))
    return out


def lot_origins(icc, nlots, seed, lperiod=True, ls=(1, 1, 1)):
    """Seeded lot origins lbeg (1-based) as four_ranloc would draw them
    (fourier_sup.f90:2513-2530), csize from four_csize (:2533-2555)."""
    rng = np.random.default_rng(seed)
    csize = np.array(icc if lperiod else [max(1, icc[i] - ls[i] + 1) for i in range(3)])
    return (np.floor(rng.random((nlots, 3)) * csize).astype(np.int32) + 1)
