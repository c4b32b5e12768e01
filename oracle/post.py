"""CPU restatement of DISCUS's post-processing of a finished Fourier run, in
numpy FP64.

TEST INFRASTRUCTURE ONLY: tests/ and bench.py's cpu_baseline leg use this
module as the checker.  The product package (diffusecode_amd) never imports
it.

Arrays are in DISCUS's layout: shape num = (n1, n2, n3), index (i, j, k) with
the abscissa fastest in memory (Fortran order); 0-based here.

What each function follows, and what pins it (tests/test_oracle_post.py):

  dsort             four_conv's MAXLOC sequence (fourier_conv.f90:18-28).
  map_fwd_* / map_rev_*
                    lib_f90/map_1dtofield.f90: maptofftfd_1D (:39-89),
                    _22D (:165-236), _33D (:325-403); fftfdtoline_1D
                    (:411-453), _22D (:556-632), _33D (:718-796).  Restated
                    literally, including the dsort permutation of the FFT
                    axes and the parity term l of each axis.
  four_conv         fourier_conv.f90:7-423: Gaussian profile through
                    diff_tr at pixel offsets, FFT, product, a second FORWARD
                    FFT, reverse map, integral kept through sum(dsi).
                    Pinned by: a brute-force direct circular convolution
                    with the profile, renormalised, equals it on odd grids
                    (1D, 2D, 3D); on even 1D lines it equals the direct
                    result advanced by one pixel (out[p] = conv[p+1]), and on
                    2D grids with num(2) > num(1) of different parity it is
                    shifted by one pixel along both axes; derived by hand
                    from the two maps and the double forward transform.
  four_filter       do_four_filter (fourier_sup.f90:1646-1805) over
                    do_lanczos_*_fft (lib_lanczos.f90:258-458), wsinc/lqgm
                    (:462-608), sinc (lib_functions.f90:10-30) and
                    do_convolute_*_real (lib_conv.f90:10-269).  Pinned by:
                    direct circular convolution with the sinc table, centred
                    at pixel (N+1)/2 (1-based), renormalised; shift rules of
                    the even-length centre derived by hand.
  four_weight       fourier_sup.f90:1428-1644 with do_volume
                    (metric.f90:576-605).  Pinned by: the voxel volume
                    through lengths and cosines equals |det(bases)| * vr for
                    the reciprocal metric (closed form), and Bragg pixels of
                    a 3D grid are divided by ncells*|det(vi)|.
  four_symm         four_symm_csf/_dsi (fourier_sup.f90:4048-4293), literal
                    scatter.  Pinned by: the output of a full point group on
                    a centred grid is invariant under every operation; a
                    constant field gives c / nops where the whole orbit is in
                    the grid; a 2-fold axis gives (x + x_rot) / 2 / 2.
  four_fill_bragg   four_fill_csf/_dsi Bragg fill for fave < 0
                    (fourier_sup.f90:3844-4046), literal sequential loop.
                    Pinned by: a single Bragg pixel becomes the mean of its
                    in-grid neighbours (corner, edge, interior cases).
  qval              qval.f90:65-286 for the values that need no random
                    numbers or PDF.  Pinned by closed forms: |F|^2, |F|,
                    phase in degrees, Re, Im, I(Q) = dsi / N, <f^2>, <f>^2
                    of a single-type crystal equal |f|^2 and |f|^2.
"""
from __future__ import annotations

import numpy as np

F32_1EM4 = float(np.float32(1.0e-4))       # DISCUS single-precision literals
F32_1EM5 = float(np.float32(1.0e-5))
F32_1EM6 = float(np.float32(1.0e-6))
F32_0005 = float(np.float32(0.005))
RAD_SP = float(np.float32(np.pi / 180.0))
ZPI = 2.0 * np.pi


def cdiv(x, r):
    """x / r for a real divisor r, component by component as Fortran does
    (numpy's complex division multiplies by 1/r and can differ in the last
    bit)."""
    x = np.asarray(x)
    if np.iscomplexobj(x):
        return (x.real / r) + 1j * (x.imag / r)
    return x / r


def nint(x):
    """Fortran NINT: round half away from zero."""
    return int(np.floor(abs(x) + 0.5) * (1 if x >= 0 else -1))


def dsort(num):
    """Axes by decreasing size, ties to the lower axis (MAXLOC)."""
    return sorted(range(3), key=lambda a: (-num[a], a))


def _par(n):
    return (n - 1) % 2


# ---------------------------------------------------------------- maps
def map_fwd_1d(line):
    n = line.shape[0]
    out = np.zeros(n, np.complex128)
    ii = (np.arange(n) + 1 + n // 2 - _par(n)) % n
    out[ii] = line
    return out


def map_rev_1d(fft):
    n = fft.shape[0]
    loop = n - 1 - ((np.arange(n) + n // 2) % n)
    out = np.zeros(n, fft.dtype)
    out[loop] = fft
    return out


def map_fwd_22d(arr, dimen, ds):
    """arr (dimen[0], dimen[1]) natural; FFT field (dimen[ds0], dimen[ds1])."""
    l = [_par(dimen[0]), _par(dimen[1])]
    ll = [(np.arange(dimen[a]) + 1 + dimen[a] // 2 - l[a]) % dimen[a] for a in range(2)]
    out = np.zeros((dimen[ds[0]], dimen[ds[1]]), np.complex128)
    I, J = np.meshgrid(ll[0], ll[1], indexing="ij")
    idx = (I, J)
    out[idx[ds[0]], idx[ds[1]]] = arr
    return out


def map_rev_22d(fft, dimen, ds):
    fd = fft.shape
    l = [_par(fd[0]), _par(fd[1]), 0]
    l[ds[0]] = _par(dimen[ds[0]])            # the overwrite of fftfdtoline_22D (:577-578)
    l[ds[1]] = _par(dimen[ds[1]])
    out = np.zeros((dimen[0], dimen[1]), fft.dtype)
    ii = np.arange(fd[0])
    jj = np.arange(fd[1])
    lla = dimen[ds[0]] - 1 - ((ii + fd[0] // 2 - l[0]) % fd[0])
    llb = dimen[ds[1]] - 1 - ((jj + fd[1] // 2 - l[1]) % fd[1])
    A, B = np.meshgrid(lla, llb, indexing="ij")
    nat = [None, None]
    nat[ds[0]] = A
    nat[ds[1]] = B
    out[nat[0], nat[1]] = fft
    return out


def map_fwd_33d(arr, dimen, ds):
    l = [_par(d) for d in dimen]
    ll = [(np.arange(dimen[a]) + 1 + dimen[a] // 2 - l[a]) % dimen[a] for a in range(3)]
    G = np.meshgrid(*ll, indexing="ij")
    out = np.zeros(tuple(dimen[a] for a in ds), np.complex128)
    out[G[ds[0]], G[ds[1]], G[ds[2]]] = arr
    return out


def map_rev_33d(fft, dimen, ds):
    fd = fft.shape
    l = [_par(f) for f in fd]
    ll = [dimen[ds[a]] - 1 - ((np.arange(fd[a]) + fd[a] // 2 - l[a]) % fd[a]) for a in range(3)]
    G = np.meshgrid(*ll, indexing="ij")
    nat = [None] * 3
    for a in range(3):
        nat[ds[a]] = G[a]
    out = np.zeros(tuple(dimen), fft.dtype)
    out[nat[0], nat[1], nat[2]] = fft
    return out


# ---------------------------------------------------------------- four_conv
def _gauss(p, s):
    return 1.0 / np.sqrt(ZPI) / s * np.exp(-0.5 * p ** 2)


def _offsets(n):
    i1 = np.arange(n)
    return np.where(i1 > n // 2, i1 - n, i1).astype(float)


def conv_profile(num, sigma, tr, isdim):
    """Profile in FFT layout = natural offsets (fourier_conv.f90:96-104, 197-212, 336-355)."""
    tr = np.asarray(tr, float).reshape(3, 3, order="F")
    if isdim == 1:
        v = _offsets(num[0])
        return _gauss(tr[0, 0] * v, sigma[0])
    if isdim == 2:
        I, J = np.meshgrid(_offsets(num[0]), _offsets(num[1]), indexing="ij")
        p1 = tr[0, 0] * I + tr[0, 1] * J
        p2 = tr[1, 0] * I + tr[1, 1] * J
        return _gauss(p1, sigma[0]) * _gauss(p2, sigma[1])
    I, J, K = np.meshgrid(_offsets(num[0]), _offsets(num[1]), _offsets(num[2]), indexing="ij")
    p = [tr[r, 0] * I + tr[r, 1] * J + tr[r, 2] * K for r in range(3)]
    return _gauss(p[0], sigma[0]) * _gauss(p[1], sigma[1]) * _gauss(p[2], sigma[2])


def four_conv(num, sigma, tr, csf, dsi, dsi3d=None, lots_off=False, fave=1.0):
    """Returns (csf, dsi, dsi3d) after four_conv; inputs are not modified."""
    num = [int(n) for n in num]
    csf = np.array(csf, np.complex128).reshape(num)
    dsi = np.array(dsi, float).reshape(num)
    d3 = None if dsi3d is None else np.array(dsi3d, float).reshape(num)
    if all(s < F32_1EM6 for s in sigma[:3]):
        return csf, dsi, d3
    ds = dsort(num)
    isdim = sum(1 for n in num if n != 1)
    prof = conv_profile(num, sigma, tr, isdim)
    ntot = num[0] * num[1] * num[2]
    if isdim == 1:
        if num[0] == 1:
            raise ValueError("four_conv_1D needs the line along the abscissa")
        n = num[0]
        temp = np.fft.fft(prof) / n

        def run(line):
            a = np.fft.fft(map_fwd_1d(line))
            a = np.fft.fft(temp * a) / float(n)
            return map_rev_1d(a)
        if lots_off:
            csf[:, 0, 0] = run(csf[:, 0, 0])
        w = dsi.sum()
        dsi[:, 0, 0] = run(dsi[:, 0, 0]).real
    elif isdim == 2:
        if num[2] != 1:
            raise ValueError("four_conv_2D needs num(3) == 1")
        prof_t = prof if ds[0] == 0 else prof.T
        temp = np.fft.fftn(prof_t) / np.sqrt(float(num[ds[0]] * num[ds[1]]))
        nn = float(np.float32(num[0] * num[1]))

        def run(a2):
            a = np.fft.fftn(map_fwd_22d(a2, num, ds))
            a = np.fft.fftn(temp * a) / nn
            return map_rev_22d(a, num, ds)
        if lots_off:
            csf[:, :, 0] = run(csf[:, :, 0])
        w = dsi.sum()
        dsi[:, :, 0] = run(dsi[:, :, 0]).real
    else:
        prof_t = np.transpose(prof, ds)
        temp = np.fft.fftn(prof_t) / np.sqrt(float(num[ds[0]] * num[ds[1]] * num[ds[2]]))
        nn = float(np.float32(ntot))

        def run(a3):
            a = np.fft.fftn(map_fwd_33d(a3, num, ds))
            a = np.fft.fftn(temp * a) / nn
            return map_rev_33d(a, num, ds)
        if lots_off:
            csf = run(csf)
        w = dsi.sum()
        dsi = run(dsi).real.copy()
    w2 = dsi.sum()
    csf = cdiv(csf * w, w2)
    dsi = dsi * w / w2
    if fave == 0.0 and d3 is not None:
        w = d3.sum()
        if isdim == 1:
            d3[:, 0, 0] = run(d3[:, 0, 0]).real
        elif isdim == 2:
            d3[:, :, 0] = run(d3[:, :, 0]).real
        else:
            d3 = run(d3).real.copy()
        d3 = d3 * w / d3.sum()
    return csf, dsi, d3


# ---------------------------------------------------------------- Lanczos
def sinc(x):
    x = np.asarray(x, float)
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(np.abs(x) < 1.0e-7, 1.0, np.sin(x) / np.where(x == 0, 1.0, x))


def lanczos_table(dims, nscale, rcut, m):
    """sinc_table of do_lanczos_*_fft over the natural indices 1..N (writes
    outside 1..N dropped)."""
    nd = len(dims)
    cen = [(n + 1) // 2 for n in dims]
    rng = []
    for a in range(nd):
        lo = nint(-nscale * m - 1.0) + cen[a]
        hi = nint(nscale * m + 1.0) + cen[a]
        i = np.arange(max(1, lo), min(dims[a], hi) + 1)
        rng.append(i)
    tab = np.zeros(dims, float)
    G = np.meshgrid(*rng, indexing="ij")
    qg = [cen[a] / nscale - G[a] / nscale for a in range(nd)]
    if nd == 1:
        w = sinc(np.pi * qg[0] * rcut) ** 2
    else:
        w = np.ones_like(qg[0])
        for a in range(nd):
            w = w * sinc(np.pi * qg[a] * rcut)
    inside = np.ones_like(qg[0], bool)
    for a in range(nd):
        inside &= np.abs(qg[a]) < float(m)
    lq = np.ones_like(qg[0])
    for a in range(nd):
        lq = lq * sinc(np.pi * qg[a] / float(m))
    lq = np.where(inside, lq, 0.0)
    idx = tuple(g - 1 for g in G)
    tab[idx] = w * lq
    return tab


def _convolute(field, table):
    nd = field.ndim
    dims = list(field.shape)
    if nd == 1:
        t = np.fft.fft(map_fwd_1d(table))
        a = np.fft.fft(map_fwd_1d(field))
        a = np.fft.fft(t * a)
        out = map_rev_1d(a).real
        return out / np.sqrt(float(dims[0]) ** 3)
    if nd == 2:
        dimen = dims + [1]
        ds = [0, 1, 2]
        t = np.fft.fftn(map_fwd_22d(table, dimen, ds))
        a = np.fft.fftn(map_fwd_22d(field, dimen, ds))
        a = np.fft.fftn(t * a)
        out = map_rev_22d(a, dimen, ds).real
        return out / (np.sqrt(float(dims[0]) ** 3) * np.sqrt(float(dims[1]) ** 3))
    ds = [0, 1, 2]
    t = np.fft.fftn(map_fwd_33d(table, dims, ds))
    a = np.fft.fftn(map_fwd_33d(field, dims, ds))
    a = np.fft.fftn(t * a)
    out = map_rev_33d(a, dims, ds).real
    return out / (np.sqrt(float(dims[0]) ** 3) * np.sqrt(float(dims[1]) ** 3) * np.sqrt(float(dims[2]) ** 3))


def lanczos(field, nscale, rcut, m):
    if nint(nscale) != 1:
        raise ValueError("output stride nint(rscale) must be 1")
    tab = lanczos_table(list(field.shape), nscale, rcut, m)
    out = _convolute(np.array(field, float), tab)
    return out * (field.sum() / out.sum() * nscale)


def four_filter(dsi, num, rscale, damp, width):
    """do_four_filter with four_filter = LANCZOS."""
    num = [int(n) for n in num]
    d = np.array(dsi, float).reshape(num)
    isd = sum(1 for n in num if n > 1)
    lim = 2 * width + 1
    if isd == 3:
        if all(n > lim for n in num):
            d = lanczos(d, rscale, damp, width)
        elif num[2] <= lim:
            for k in range(num[2]):
                d[:, :, k] = lanczos(d[:, :, k], rscale, damp, width)
        elif num[1] <= lim:
            for j in range(num[1]):
                d[:, j, :] = lanczos(d[:, j, :], rscale, damp, width)
        elif num[0] <= lim:
            for i in range(num[0]):
                d[i, :, :] = lanczos(d[i, :, :], rscale, damp, width)
    elif isd == 2:
        if num[2] != 1:
            raise ValueError("2D filter needs num(3) == 1")
        d[:, :, 0] = lanczos(d[:, :, 0], rscale, damp, width)
    elif isd == 1:
        if num[0] == 1:
            raise ValueError("1D filter needs the line along the abscissa")
        d[:, 0, 0] = lanczos(d[:, 0, 0], rscale, damp, width)
    return d


# ---------------------------------------------------------------- four_weight
def do_volume_recip(u, v, w, rten):
    """do_volume(.false., u, v, w) through lengths and cosines."""
    g = np.asarray(rten, float).reshape(3, 3, order="F")

    def blen(a):
        return np.sqrt(a @ g @ a)

    def cosang(a, b):
        # cosd(do_bang(...)): the angle comes back in degrees through a
        # single-precision rad (metric.f90:70-74)
        arg = (a @ g @ b) / (blen(a) * blen(b))
        if abs(arg) < 1.0:
            deg = np.arccos(arg) / RAD_SP
        elif arg >= 1.0:
            deg = 0.0
        else:
            deg = 180.0
        return np.cos(deg * (np.pi / 180.0))
    L = [blen(np.asarray(x, float)) for x in (u, v, w)]
    c1, c2, c3 = cosang(v, w), cosang(u, w), cosang(u, v)
    return L[0] * L[1] * L[2] * np.sqrt(1.0 - c1 ** 2 - c2 ** 2 - c3 ** 2 + 2.0 * c1 * c2 * c3)


def weight_bases(num, vi, nlots, icc, ls):
    vi = np.asarray(vi, float).reshape(3, 3, order="F")
    uin, vin, win = vi[:, 0], vi[:, 1], vi[:, 2]
    isdim = sum(1 for n in num if n != 1)
    b = np.zeros((3, 3))
    cells = icc if nlots == 1 else ls
    if isdim == 1:
        if num[0] != 1:
            b[:, 0] = uin
            ax, col = 0, 0
        elif num[1] == 1:
            b[:, 1] = vin
            ax, col = 1, 1
        else:
            b[:, 2] = win
            ax, col = 2, 2
        # the two unit vectors completing the basis (fourier_sup.f90:1492-1540)
        others = {0: (1, 2), 1: (0, 2), 2: (0, 1)}[col]
        if b[0, col] != 0.0:
            b[1, others[0]] = 1.0
            b[2, others[1]] = 1.0
        elif b[1, col] != 0.0:
            b[2, others[0]] = 1.0
            b[0, others[1]] = 1.0
        elif b[2, col] != 0.0:
            b[0, others[0]] = 1.0
            b[1, others[1]] = 1.0
        ncells = cells[ax]
    elif isdim == 2:
        if num[2] == 1:
            b[:, 0], b[:, 1], b[:, 2] = uin, vin, 1.0
            for r in range(3):
                if b[r, 0] != 0.0 or b[r, 1] != 0.0:
                    b[r, 2] = 0.0
            ncells = cells[0] * cells[1]
        elif num[1] == 1:
            b[:, 0], b[:, 1], b[:, 2] = uin, 1.0, win
            for r in range(3):
                if b[r, 0] != 0.0 or b[r, 2] != 0.0:
                    b[r, 1] = 0.0
            ncells = cells[0] * cells[2]
        else:
            raise ValueError("four_weight leaves the bases undefined for a 2D grid with num(1) == 1")
    elif isdim == 3:
        b[:, 0], b[:, 1], b[:, 2] = uin, vin, win
        ncells = cells[0] * cells[1] * cells[2]
    else:
        raise ValueError("four_weight leaves the bases undefined for a single point")
    return b, ncells


def four_weight(dsi, num, eck, vi, rten, vr, fave, nlots, icc, ls):
    """dsi3d of four_weight; dsi is the intensity before diff_scale."""
    num = [int(n) for n in num]
    d = np.array(dsi, float).reshape(num)
    scalef = 1.0
    if fave == 0.0:
        b, ncells = weight_bases(num, vi, nlots, icc, ls)
        vox = do_volume_recip(b[:, 0], b[:, 1], b[:, 2], rten)
        if vox > 0.0:
            scalef = float(ncells) / vr * vox
    d3 = d.copy()
    if fave == 0.0 and abs(scalef - 1.0) > F32_1EM5:
        e = np.asarray(eck, float).reshape(3, 4, order="F")[:, 0]
        v = np.asarray(vi, float).reshape(3, 3, order="F")
        I, J, K = np.meshgrid(*[np.arange(n, dtype=float) for n in num], indexing="ij")
        ok = np.ones(num, bool)
        for r in range(3):
            h = e[r] + v[r, 0] * I + v[r, 1] * J + v[r, 2] * K
            ok &= np.abs(h - np.where(h >= 0, np.floor(h + 0.5), -np.floor(-h + 0.5))) < F32_1EM4
        d3[ok] = d[ok] / scalef
    return d3, scalef


# ---------------------------------------------------------------- symmetry
def symm_izero(num, eck, vi, unit=1.0):
    e = np.asarray(eck, float).reshape(3, 4, order="F")[:, 0]
    v = np.asarray(vi, float).reshape(3, 3, order="F")
    m = np.zeros((3, 3))
    for a in range(3):
        if num[a] > 1:
            m[:, a] = v[:, a]
        else:
            m[a, a] = unit
    uvw = np.linalg.inv(m) @ (-e)
    return [nint(1.0 + uvw[a]) - 1 if num[a] > 1 else 0 for a in range(3)]


def four_symm(arr, num, eck, vi, mats, nops, unit=1.0):
    """mats: (3, 3, nops-1) Fortran order, operations 2..nops of the point
    group part; nops = spc_n / n_center.  Literal scatter."""
    num = [int(n) for n in num]
    x = np.array(arr).reshape(num)
    iz = symm_izero(num, eck, vi, unit)
    sym = x.copy()
    wgt = np.ones(num, np.int64)
    M = np.asarray(mats, float).reshape(3, 3, -1, order="F")
    I, J, K = np.meshgrid(*[np.arange(n) for n in num], indexing="ij")
    hkl = np.stack([(I - iz[0]).ravel(order="F"), (J - iz[1]).ravel(order="F"), (K - iz[2]).ravel(order="F")], 1).astype(float)
    src = x.ravel(order="F")
    for op in range(M.shape[2]):
        uvw = hkl @ M[:, :, op]                   # MATMUL(hkl, sym_mat)
        t = np.stack([iz[a] + np.vectorize(nint)(uvw[:, a]) for a in range(3)], 1) if uvw.size else np.zeros((0, 3), int)
        ok = np.all((t >= 0) & (t < np.asarray(num)), axis=1)
        for p in np.nonzero(ok)[0]:               # sequential, in DISCUS's loop order
            ti = tuple(t[p])
            sym[ti] = sym[ti] + src[p]
            wgt[ti] += 1
    return cdiv(cdiv(sym, wgt), float(nops))


def four_fill_bragg(arr, num, eck, vi):
    """Bragg fill of four_fill_csf/_dsi for fave < 0, in DISCUS's loop order."""
    num = [int(n) for n in num]
    x = np.array(arr).reshape(num).copy()
    e = np.asarray(eck, float).reshape(3, 4, order="F")[:, 0]
    v = np.asarray(vi, float).reshape(3, 3, order="F")
    step = [max(v[r, 0], v[r, 1], v[r, 2]) * F32_0005 for r in range(3)]
    for a in range(3):
        if num[a] == 1:
            step[a] = float(np.float32(0.0001))
    for h in range(num[0]):
        for k in range(num[1]):
            for l in range(num[2]):
                hkl = e + v[:, 0] * h + v[:, 1] * k + v[:, 2] * l
                if all(abs(hkl[r] - nint(hkl[r])) < step[r] for r in range(3)):
                    acc = 0.0
                    n = 0
                    for (a, b, c, ok) in ((h - 1, k, l, h > 0), (h + 1, k, l, h < num[0] - 1),
                                          (h, k - 1, l, k > 0), (h, k + 1, l, k < num[1] - 1),
                                          (h, k, l - 1, l > 0), (h, k, l + 1, l < num[2] - 1)):
                        if ok:
                            acc = acc + x[a, b, c]
                            n += 1
                    if n > 0:
                        x[h, k, l] = cdiv(acc, float(n))
    return x


# ---------------------------------------------------------------- qval
VAL_INTEN, VAL_AMPLI, VAL_PHASE, VAL_REAL, VAL_IMAG = 1, 2, 3, 4, 5
VAL_SQ, VAL_FQ, VAL_F2AVER, VAL_FAVER2, VAL_FAVER = 7, 8, 9, 10, 11
VAL_IQ, VAL_F2AVERB, VAL_FAVER2B, VAL_FAVERB = 13, 17, 18, 19


def qval(value, laver, csf, acsf, dsi, istl, cfact, cfact_pure, amount, n_real, u2aver, cfinc,
         faver_sign=1.0):
    """Per-pixel qval over the whole grid (arrays of shape num).  istl is
    0-based into cfact rows as DISCUS's 0:CFPKT; cfact/_pure (ncfact, nscat)."""
    f = np.asarray(acsf if laver else csf)
    d = np.asarray(dsi, float)
    n1, n2, n3 = f.shape
    if value == VAL_INTEN:
        return (f * np.conj(f)).real if laver else d.copy()
    if value == VAL_AMPLI:
        return np.sqrt((f * np.conj(f)).real)
    if value == VAL_PHASE:
        return np.where(f == 0, 0.0, np.degrees(np.arctan2(f.imag, f.real)))
    if value == VAL_REAL:
        return f.real.copy()
    if value == VAL_IMAG:
        return f.imag.copy()
    it = np.asarray(istl)
    ij1 = np.repeat(it[:, :, :1], n3, axis=2)        # istl(i, j, 1)
    i11 = np.broadcast_to(it[:, :1, :1], it.shape)   # istl(i, 1, 1)
    cf = np.asarray(cfact)
    cp = np.asarray(cfact_pure)
    nsc = cf.shape[1]
    amt = np.asarray(amount, float)

    def sum_abs(tab, signed, sign_fixed=None):
        s = np.zeros(f.shape)
        for t in range(nsc):
            sg = 1.0
            if signed:
                sg = -1.0 if cp[1, t].real < 0.0 else 1.0      # cfact_pure(1, l)
            if sign_fixed is not None:
                sg = sign_fixed
            s = s + np.sqrt((tab[ij1, t] * np.conj(tab[ij1, t])).real) * amt[t] * sg
        return s

    def sum_sq(tab):
        s = np.zeros(f.shape)
        for t in range(nsc):
            s = s + (tab[ij1, t] * np.conj(tab[ij1, t])).real * amt[t]
        return s
    if value == VAL_SQ:
        q = (f * np.conj(f)).real if laver else d.copy()
        fa2 = sum_abs(cp, True) ** 2
        q2 = (ZPI ** 2 * (2.0 * i11 * cfinc) ** 2).astype(np.float32).astype(float)   # REAL() to default kind
        return q / fa2 / n_real + 1.0 - np.exp(-q2 * u2aver)
    if value == VAL_IQ:
        q = (f * np.conj(f)).real if laver else d.copy()
        return q / n_real
    if value == VAL_F2AVER:
        return sum_sq(cp) / n_real
    if value == VAL_FAVER2:
        return (sum_abs(cp, True) / n_real) ** 2
    if value == VAL_FAVER:
        return sum_abs(cp, False, faver_sign) / n_real
    if value == VAL_F2AVERB:
        return sum_sq(cf) / n_real
    if value == VAL_FAVER2B:
        return (sum_abs(cf, True) / n_real) ** 2
    if value == VAL_FAVERB:
        return sum_abs(cf, False, faver_sign) / n_real
    raise ValueError(f"qval value {value} not restated")
