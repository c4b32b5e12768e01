"""Pins of the post-processing oracle (oracle/post.py) by checks that do not
retype its formulas: direct circular convolutions (brute force), shift rules
derived by hand from the FFT-layout maps, closed-form volumes and group
invariance."""
import numpy as np
import pytest

from oracle import post as P


def direct_circ(x, kern_of_offset):
    """out[p] = sum_q x[q] k(p - q), offsets wrapped into (-N/2, N/2]."""
    shp = x.shape
    out = np.zeros(shp, np.result_type(x, float))
    grids = np.meshgrid(*[np.arange(n) for n in shp], indexing="ij")
    for q in np.ndindex(*shp):
        off = []
        for a, n in enumerate(shp):
            d = (grids[a] - q[a]) % n
            off.append(np.where(d > n // 2, d - n, d))
        out = out + x[q] * kern_of_offset(*off)
    return out


def gauss_kernel(sigma, tr):
    tr = np.asarray(tr, float).reshape(3, 3, order="F")

    def k(*off):
        v = list(off) + [0] * (3 - len(off))
        val = 1.0
        for r in range(len(off)):
            p = tr[r, 0] * v[0] + tr[r, 1] * v[1] + tr[r, 2] * v[2]
            val = val * np.exp(-0.5 * p ** 2) / np.sqrt(2 * np.pi) / sigma[r]
        return val
    return k


def renorm(conv_dsi, dsi):
    return conv_dsi * dsi.sum() / conv_dsi.sum()


TR = [0.7, 0.1, 0.0, -0.05, 0.9, 0.0, 0.0, 0.02, 1.1]
SIG = [1.0, 1.3, 0.8]


@pytest.mark.parametrize("n", [9, 10])
def test_conv_1d_direct(n):
    rng = np.random.default_rng(n)
    dsi = rng.random(n)
    csf = rng.random(n) + 1j * rng.random(n)
    c, d, _ = P.four_conv((n, 1, 1), SIG, TR, csf.reshape(n, 1, 1), dsi.reshape(n, 1, 1), lots_off=True)
    k = gauss_kernel(SIG, TR)
    cd = direct_circ(dsi, lambda o: k(o))
    cc = direct_circ(csf, lambda o: k(o))
    if n % 2 == 0:                      # even line: out[p] = conv[p + 1]
        cd, cc = np.roll(cd, -1), np.roll(cc, -1)
    want_d = renorm(cd, dsi)
    want_c = cc * dsi.sum() / cd.sum()
    assert np.allclose(d[:, 0, 0], want_d, rtol=1e-12, atol=1e-14)
    assert np.allclose(c[:, 0, 0], want_c, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("shape,shift", [((7, 5), (0, 0)), ((6, 8), (0, 0)), ((5, 8), (1, -1)), ((6, 9), (-1, 1))])
def test_conv_2d_direct(shape, shift):
    """shift: where the reverse map of a transposed field (num(2) > num(1))
    uses the other axis's parity: out[p] = conv[p - shift] per axis, from
    ll0 = p - (l - l') of the reverse map."""
    n1, n2 = shape
    rng = np.random.default_rng(n1 * 10 + n2)
    dsi = rng.random(shape)
    csf = rng.random(shape) + 1j * rng.random(shape)
    c, d, _ = P.four_conv((n1, n2, 1), SIG, TR, csf[:, :, None], dsi[:, :, None], lots_off=True)
    k = gauss_kernel(SIG, TR)
    cd = direct_circ(dsi, lambda a, b: k(a, b))
    cc = direct_circ(csf, lambda a, b: k(a, b))
    cd = np.roll(cd, shift, axis=(0, 1))
    cc = np.roll(cc, shift, axis=(0, 1))
    assert np.allclose(d[:, :, 0], renorm(cd, dsi), rtol=1e-12, atol=1e-14)
    assert np.allclose(c[:, :, 0], cc * dsi.sum() / cd.sum(), rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("shape", [(5, 7, 3), (4, 6, 5), (3, 5, 8)])
def test_conv_3d_direct(shape):
    rng = np.random.default_rng(sum(shape))
    dsi = rng.random(shape)
    d3 = rng.random(shape)
    c, d, e = P.four_conv(shape, SIG, TR, np.zeros(shape, complex), dsi, d3, lots_off=False, fave=0.0)
    k = gauss_kernel(SIG, TR)
    assert np.allclose(d, renorm(direct_circ(dsi, k), dsi), rtol=1e-12, atol=1e-14)
    assert np.allclose(e, renorm(direct_circ(d3, k), d3), rtol=1e-12, atol=1e-14)


def test_conv_off_below_eps():
    x = np.arange(6.0).reshape(6, 1, 1)
    c, d, _ = P.four_conv((6, 1, 1), [5e-7, 5e-7, 5e-7], TR, x.astype(complex), x)
    assert np.array_equal(d, x)


def sinc(x):
    return np.where(np.abs(x) < 1e-7, 1.0, np.sin(x) / np.where(x == 0, 1, x))


def lanczos_kernel(nd, rcut, m, nscale=1.0):
    def k(*off):
        t = [np.asarray(o, float) / nscale for o in off]
        inside = np.ones_like(t[0], bool)
        lq = np.ones_like(t[0])
        w = np.ones_like(t[0])
        for x in t:
            inside &= np.abs(x) < m
            lq = lq * sinc(np.pi * x / m)
            w = w * sinc(np.pi * x * rcut)
        if nd == 1:
            w = w * w
        return np.where(inside, w * lq, 0.0)
    return k


@pytest.mark.parametrize("n", [21, 22])
def test_lanczos_1d_direct(n):
    """Odd: plain convolution.  Even: the table centre sits at FFT index -1
    (shift 1) and the 1D reverse map adds one more (shift 1): out[p] = conv[p+2]."""
    rng = np.random.default_rng(n)
    x = rng.random(n)
    out = P.lanczos(x, 1.0, 0.5, 4)
    cd = direct_circ(x, lanczos_kernel(1, 0.5, 4))
    if n % 2 == 0:
        cd = np.roll(cd, -2)
    assert np.allclose(out, cd * x.sum() / cd.sum(), rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("shape", [(11, 13), (12, 11), (11, 11, 13), (12, 12, 11)])
def test_lanczos_nd_direct(shape):
    """2D/3D: the maps are exact inverses; an even axis shifts by one (the
    table centre at FFT index -1): out[p] = conv[p+1] along even axes."""
    rng = np.random.default_rng(sum(shape))
    x = rng.random(shape)
    out = P.lanczos(x, 1.0, 0.5, 4)
    cd = direct_circ(x, lanczos_kernel(len(shape), 0.5, 4))
    for a, n in enumerate(shape):
        if n % 2 == 0:
            cd = np.roll(cd, -1, axis=a)
    assert np.allclose(out, cd * x.sum() / cd.sum(), rtol=1e-12, atol=1e-14)


def test_filter_dispatch_slices():
    """A 3D grid thinner than 2*width+1 along the top axis is filtered as
    independent 2D layers."""
    rng = np.random.default_rng(7)
    x = rng.random((11, 13, 3))
    out = P.four_filter(x, (11, 13, 3), 1.0, 0.5, 4)
    for k in range(3):
        assert np.allclose(out[:, :, k], P.lanczos(x[:, :, k], 1.0, 0.5, 4), rtol=1e-14)


def test_volume_closed_form():
    a, b, c = 5.0, 6.0, 7.0
    al, be, ga = np.radians([80.0, 95.0, 110.0])
    g = np.array([[a * a, a * b * np.cos(ga), a * c * np.cos(be)],
                  [a * b * np.cos(ga), b * b, b * c * np.cos(al)],
                  [a * c * np.cos(be), b * c * np.cos(al), c * c]])
    rten = np.linalg.inv(g)
    vr = np.sqrt(np.linalg.det(rten))
    B = np.array([[0.1, 0.02, 0.0], [0.0, 0.05, 0.01], [0.03, 0.0, 0.2]])
    vol = P.do_volume_recip(B[:, 0], B[:, 1], B[:, 2], rten.ravel(order="F"))
    assert vol == pytest.approx(abs(np.linalg.det(B)) * vr, rel=1e-7)


def test_weight_bragg_3d():
    rten = np.diag([1 / 25.0, 1 / 25.0, 1 / 25.0])
    vr = 1 / 125.0
    num = (9, 9, 5)
    vi = np.diag([0.25, 0.25, 0.5]).ravel(order="F")
    eck = np.zeros((3, 4))
    eck[:, 0] = [-1.0, -1.0, -1.0]
    rng = np.random.default_rng(3)
    dsi = rng.random(num)
    d3, scalef = P.four_weight(dsi, num, eck.ravel(order="F"), vi, rten.ravel(order="F"), vr, 0.0, 1, (4, 4, 4), (2, 2, 2))
    assert scalef == pytest.approx(64 * 0.25 * 0.25 * 0.5, rel=1e-7)
    bragg = np.zeros(num, bool)
    bragg[0::4, 0::4, 0::2] = True
    assert np.allclose(d3[bragg], dsi[bragg] / scalef, rtol=1e-15)
    assert np.array_equal(d3[~bragg], dsi[~bragg])


def test_weight_fave_leaves_dsi():
    x = np.arange(12.0).reshape(4, 3, 1)
    d3, s = P.four_weight(x, (4, 3, 1), np.zeros(12), np.eye(3).ravel(), np.eye(3).ravel(), 1.0, 50.0, 1,
                          (4, 4, 4), (2, 2, 2))
    assert s == 1.0 and np.array_equal(d3, x)


def _centred(n):
    eck = np.zeros((3, 4))
    eck[:, 0] = [-(n[0] // 2) * 0.1, -(n[1] // 2) * 0.1, -(n[2] // 2) * 0.1 if n[2] > 1 else 0.0]
    return eck.ravel(order="F"), (np.eye(3) * 0.1).ravel(order="F")


def c4v_ops():
    r4 = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], float)
    mx = np.diag([-1.0, 1, 1])
    ops = []
    g = np.eye(3)
    for _ in range(4):
        ops.append(g.copy())
        ops.append(g @ mx)
        g = g @ r4
    return ops                                       # 8 operations, identity first


def test_symm_group_invariance():
    num = (9, 9, 1)
    eck, vi = _centred(num)
    rng = np.random.default_rng(1)
    x = rng.random(num)
    ops = c4v_ops()
    mats = np.stack(ops[1:], axis=2).ravel(order="F")
    out = P.four_symm(x, num, eck, vi, mats, 8)
    c = 4
    for S in ops:
        for i in range(9):
            for j in range(9):
                t = np.array([i - c, j - c, 0]) @ S
                assert out[i, j, 0] == pytest.approx(out[int(t[0]) + c, int(t[1]) + c, 0], rel=1e-13)
    # the whole orbit is inside the grid: weight = 8 everywhere
    const = P.four_symm(np.full(num, 3.0), num, eck, vi, mats, 8)
    assert np.allclose(const, 3.0 / 8.0)


def test_symm_twofold():
    num = (7, 5, 1)
    eck, vi = _centred(num)
    rng = np.random.default_rng(2)
    x = rng.random(num) + 1j * rng.random(num)
    two = np.diag([-1.0, -1.0, 1.0])
    out = P.four_symm(x, num, eck, vi, two.ravel(order="F"), 2)
    assert np.allclose(out[:, :, 0], (x[:, :, 0] + x[::-1, ::-1, 0]) / 2 / 2, rtol=1e-15)


@pytest.mark.parametrize("pos", [(0, 0, 0), (4, 0, 2), (4, 3, 2)])
def test_fill_single_bragg(pos):
    num = (9, 7, 5)
    vi = (np.eye(3) * 0.3).ravel(order="F")
    eck = np.zeros((3, 4))
    eck[:, 0] = [-p * 0.3 for p in pos]
    # make only `pos` integer: steps 0.3 reach an integer only at offset 0 within the grid (|offset| < 10)
    rng = np.random.default_rng(5)
    x = rng.random(num)
    out = P.four_fill_bragg(x, num, eck.ravel(order="F"), vi)
    nb = []
    for a in range(3):
        for d in (-1, 1):
            q = list(pos)
            q[a] += d
            if 0 <= q[a] < num[a]:
                nb.append(x[tuple(q)])
    want = x.copy()
    want[pos] = np.mean(nb)
    assert np.allclose(out, want, rtol=1e-15)


def test_qval_closed_forms():
    rng = np.random.default_rng(9)
    num = (4, 3, 2)
    f = rng.random(num) + 1j * rng.random(num) - 0.5
    a = rng.random(num) + 1j * rng.random(num)
    d = rng.random(num)
    istl = rng.integers(0, 20, size=num)
    cf = rng.random((21, 1)) + 0.1j
    cp = cf * 1.5
    amt = [40.0]
    args = (f, a, d, istl, cf, cp, amt, 40.0, 0.01, 0.001)
    assert np.allclose(P.qval(P.VAL_INTEN, False, *args), d)
    assert np.allclose(P.qval(P.VAL_INTEN, True, *args), np.abs(a) ** 2)
    assert np.allclose(P.qval(P.VAL_AMPLI, False, *args), np.abs(f))
    assert np.allclose(P.qval(P.VAL_PHASE, False, *args), np.angle(f, deg=True))
    assert np.allclose(P.qval(P.VAL_REAL, True, *args), a.real)
    assert np.allclose(P.qval(P.VAL_IMAG, False, *args), f.imag)
    assert np.allclose(P.qval(P.VAL_IQ, False, *args), d / 40.0)
    i1 = np.repeat(istl[:, :, :1], 2, axis=2)
    assert np.allclose(P.qval(P.VAL_F2AVER, False, *args), np.abs(cp[i1, 0]) ** 2)
    assert np.allclose(P.qval(P.VAL_FAVER2, False, *args), np.abs(cp[i1, 0]) ** 2)
    assert np.allclose(P.qval(P.VAL_FAVERB, False, *args), np.abs(cf[i1, 0]))


# ---------------------------------------------------------------- NUFFT checker
from oracle import nufft as NU


def test_nufft_direct_single_point():
    x = np.array([0.37])
    f = NU.nufft1_direct([x], [2.0 + 0j], 1, [7])
    k = np.arange(7) - 3
    assert np.allclose(f, 2.0 * np.exp(1j * k * 0.37), rtol=1e-15)


def test_nufft_direct_lattice_dirichlet():
    L = 8
    m = np.arange(L)
    x = 2 * np.pi * m / L * 0.3          # points 2 pi (0.3 m / L)
    f = NU.nufft1_direct([x], np.ones(L, complex), 1, [9])
    k = np.arange(9) - 4
    t = 0.3 * k / L
    want = np.where(np.abs(np.sin(np.pi * t)) < 1e-14, L,
                    np.exp(1j * np.pi * t * (L - 1)) * np.sin(np.pi * L * t) / np.where(np.sin(np.pi * t) == 0, 1, np.sin(np.pi * t)))
    assert np.allclose(f, want, rtol=1e-12, atol=1e-12)


def test_nufft_direct_conjugate_and_bruteforce():
    import cmath
    rng = np.random.default_rng(3)
    pts = [rng.uniform(-np.pi, np.pi, 3) for _ in range(3)]
    c = rng.random(3) + 0j
    fp = NU.nufft1_direct(pts, c, 1, [4, 3, 5])
    fm = NU.nufft1_direct(pts, c, -1, [4, 3, 5])
    assert np.allclose(fp, np.conj(fm), rtol=1e-13)                # real weights
    assert np.allclose(fp[1:, :, :][::-1, ::-1, ::-1], fm[1:, :, :], rtol=1e-13)   # f+(-k) = f-(k)
    for (a, b, cc) in [(0, 0, 0), (3, 2, 4), (1, 1, 2)]:
        k = (a - 2, b - 1, cc - 2)
        s = sum(c[j] * cmath.exp(1j * (k[0] * pts[0][j] + k[1] * pts[1][j] + k[2] * pts[2][j])) for j in range(3))
        assert abs(fp[a, b, cc] - s) < 1e-14


def test_strucf_nufft_is_structure_factor():
    """positions 2 pi s x / n (four_getatm_anis_iam) -> F(h = k s / n)."""
    rng = np.random.default_rng(4)
    frac = rng.random((50, 3)) * 6.0
    s, n = 2.0, 6.0
    pts = [2 * np.pi * s * frac[:, d] / n for d in range(3)]
    inc, ls = (5, 4, 3), (2, 1, 3)
    t = NU.strucf_nufft(3, pts, 0.5, inc, ls)
    for idx in [(0, 0, 0), (4, 3, 2), (2, 1, 1)]:
        k = [idx[d] * ls[d] - ((inc[d] - 1) * ls[d] + 1) // 2 for d in range(3)]
        h = np.array(k) * s / n
        want = 0.5 * np.exp(2j * np.pi * frac @ h).sum()
        assert abs(t[idx] - want) < 1e-11
