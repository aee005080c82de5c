"""Closed-form and brute-force checks of the oracle's building blocks, each
chosen so that a dropped term, a sign or index error, or a transposed
operand fails: Lagrange basis (interpolation property, partition of unity,
polynomial reproduction against scipy), DFT (numpy.fft), M2L kernel tables
(circulant embedding equals the dense Toeplitz product), symmetry
permutations (permuted fundamental table equals the table built directly for
the offset), direction cones, M2M matrices and the direct-sum kernel."""
import ctypes

import numpy as np
import pytest
from scipy.interpolate import lagrange as sp_lagrange

import oracle_api as oa

D = ctypes.POINTER(ctypes.c_double)


def p(a):
    return a.ctypes.data_as(D)


@pytest.fixture(scope="module")
def lib():
    return oa.load()


def grid(L):
    return -1.0 + 2.0 * np.arange(L) / (L - 1)


@pytest.mark.parametrize("L", [2, 3, 5, 7, 8])
def test_lagrange_basis(lib, L):
    xi = grid(L)
    for a in range(L):
        for b in range(L):
            assert abs(lib.oracle_S1D(a, 0.37 * xi[b], L, 0.37) - (a == b)) < 1e-13
    xs = np.linspace(-1.3, 1.3, 17)
    for x in xs:
        assert abs(sum(lib.oracle_S1D(a, x, L, 1.0) for a in range(L)) - 1.0) < 1e-11
    rng = np.random.default_rng(L)
    c = rng.standard_normal(L)
    poly = np.polynomial.Polynomial(c)
    for x in xs:
        v = sum(lib.oracle_S1D(a, x, L, 1.0) * poly(xi[a]) for a in range(L))
        assert abs(v - poly(x)) < 1e-9 * (1 + abs(poly(x)))
        # independent: scipy's Lagrange polynomial through the same nodes
        e = np.zeros(L)
        e[1] = 1.0
        assert abs(lib.oracle_S1D(1, x, L, 1.0) - sp_lagrange(xi, e)(x)) < 1e-8


@pytest.mark.parametrize("P", [3, 7, 9, 11, 13])
def test_dft_matches_numpy(lib, P):
    rng = np.random.default_rng(P)
    a = rng.standard_normal((P, P, P)) + 1j * rng.standard_normal((P, P, P))
    out = np.zeros_like(a)
    lib.oracle_dft3(P, p(a.view(np.float64)), p(out.view(np.float64)), 1)
    assert np.abs(out - np.fft.fftn(a)).max() < 1e-11 * np.abs(a).sum() / P
    lib.oracle_dft3(P, p(a.view(np.float64)), p(out.view(np.float64)), 0)
    assert np.abs(out - np.fft.ifftn(a) * P ** 3).max() < 1e-11 * np.abs(a).sum() / P


def _table(lib, L, kappa, rad, v):
    P = 2 * L - 1
    out = np.zeros(P ** 3, dtype=np.complex128)
    vv = np.asarray(v, dtype=np.float64)
    lib.oracle_m2l_table(L, kappa, rad, p(vv), p(out.view(np.float64)))
    return out


@pytest.mark.parametrize("L,kappa", [(3, 2.0), (5, 30.0), (6, 0.5)])
def test_m2l_table_is_toeplitz_product(lib, L, kappa):
    """IDFT(U .* DFT(embed M))/PP on the L^3 corner == sum_J G(|v + r(xi_I - xi_J)|) M_J."""
    P = 2 * L - 1
    rad = 0.125
    v = np.array([3.0, 1.0, -2.0]) * 2 * rad
    U = _table(lib, L, kappa, rad, v).reshape(P, P, P)
    rng = np.random.default_rng(1)
    M = rng.standard_normal((L, L, L)) + 1j * rng.standard_normal((L, L, L))
    E = np.zeros((P, P, P), dtype=complex)
    E[:L, :L, :L] = M
    got = np.fft.ifftn(U * np.fft.fftn(E))[:L, :L, :L]
    xi = grid(L)
    X = np.stack(np.meshgrid(xi, xi, xi, indexing="ij"), -1).reshape(-1, 3) * rad
    R = np.linalg.norm(v + X[:, None, :] - X[None, :, :], axis=-1)
    G = np.exp(1j * kappa * R) / (4 * np.pi * R)
    want = (G @ M.reshape(-1)).reshape(L, L, L)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


def _fundamental(X):
    a = np.abs(X)
    h1 = (a[0] < a[1]) | ((a[0] < a[2]) << 1) | ((a[1] < a[2]) << 2)
    h0 = (X[0] < 0) | ((X[1] < 0) << 1) | ((X[2] < 0) << 2)
    return np.array(sorted(a, reverse=True)), int(h0) + (int(h1) << 3)


@pytest.mark.parametrize("X", [(3, -1, 2), (-2, -2, 5), (0, 4, -1), (-3, 0, 0), (1, 1, -4), (2, -5, 3)])
def test_symmetry_permutation(lib, X):
    L, kappa, rad = 4, 7.0, 0.2
    P = 2 * L - 1
    X = np.array(X)
    nI, h = _fundamental(X)
    Uf = _table(lib, L, kappa, rad, 2 * rad * nI)
    Ud = _table(lib, L, kappa, rad, 2 * rad * X.astype(float))
    perm = np.zeros(P ** 3, dtype=np.int32)
    lib.oracle_perm(L, h, perm.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    assert np.abs(Uf[perm] - Ud).max() < 1e-12 * np.abs(Ud).max()


def test_directions(lib):
    n = 5
    for st in range(n):
        out = np.zeros(3 * 6 * 4 ** st)
        cnt = lib.oracle_directions(n, st, p(out))
        d = out.reshape(-1, 3)
        assert cnt == 6 * 4 ** st
        assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-14)
        assert len(np.unique(np.round(d, 12), axis=0)) == cnt
        # the set is symmetric under x -> -x
        assert len(np.unique(np.round(np.concatenate([d, -d]), 12), axis=0)) == cnt
    # stage 0: +x, -x, +y, -y, +z, -z ; stage 1 starts with the +x face
    out = np.zeros(18)
    lib.oracle_directions(n, 0, p(out))
    assert np.array_equal(out.reshape(6, 3), np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0],
                                                        [0, 0, 1], [0, 0, -1]], dtype=float))
    out = np.zeros(72)
    lib.oracle_directions(n, 1, p(out))
    d1 = out.reshape(24, 3)
    want = np.array([1.0, 0.25, 0.25]) / np.linalg.norm([1.0, 0.25, 0.25])
    assert np.allclose(d1[0], want)
    assert np.all(d1[:4, 0] > 0.9) and np.all(d1[4:8, 0] < -0.9)   # +x then -x face
    assert np.all(d1[8:12, 2] > 0.9) and np.all(d1[16:20, 1] > 0.9)  # then z, then y


@pytest.mark.parametrize("L", [3, 5, 7])
def test_m2m_matrices(lib, L):
    """A^-[a,b] = S_a((xi_b - 1)/2): the parent weight of a unit charge at the
    child's grid point b (child on the negative side), so A^- applied to a
    child expansion equals P2M of the same charges in the parent cell."""
    out = np.zeros(2 * L * L)
    lib.oracle_m2m_1d(L, p(out))
    Am, Ap = out[:L * L].reshape(L, L), out[L * L:].reshape(L, L)
    xi = grid(L)
    for a in range(L):
        e = np.zeros(L)
        e[a] = 1.0
        Sa = sp_lagrange(xi, e)
        for b in range(L):
            assert abs(Am[a, b] - Sa((xi[b] - 1) / 2)) < 1e-9
            assert abs(Ap[a, b] - Sa((xi[b] + 1) / 2)) < 1e-9
    # column sums are 1 (partition of unity at every child point)
    assert np.allclose(Am.sum(0), 1.0) and np.allclose(Ap.sum(0), 1.0)


def test_direct_kernel(lib):
    rng = np.random.default_rng(3)
    src = rng.random((40, 3))
    trg = np.concatenate([rng.random((10, 3)), src[:3]])   # includes coincident points
    q = rng.standard_normal(40) + 1j * rng.standard_normal(40)
    kappa = 9.0
    got = oa.direct(src, q, trg, kappa)
    R = np.linalg.norm(trg[:, None] - src[None], axis=-1)
    with np.errstate(divide="ignore", invalid="ignore"):
        G = np.where(R > 1e-8, np.exp(1j * kappa * R) / (4 * np.pi * R), 0.0)
    assert np.abs(got - G @ q).max() < 1e-13 * np.abs(G @ q).max()


def test_small_fmm_converges(lib):
    """Error against the direct sum falls with the interpolation order."""
    rng = np.random.default_rng(5)
    src = rng.random((3000, 3))
    q = rng.standard_normal(3000) + 1j * rng.standard_normal(3000)
    ex = oa.direct(src, q, src, 20.0)
    errs = []
    for L in (3, 5, 7):
        pf, _, _, _ = oa.fmm(src, src, q, L, 20, 20.0)
        errs.append(oa.rel_errors(ex, pf)[2])
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 1e-4
