"""Pins of the oracle's hex8 element and assembly against closed forms."""
import numpy as np
import pytest

from oracle import fe
from oracle.mesh import Mesh, opts


def brick(hx, hy, hz, x0=(0.0, 0.0, 0.0)):
    # node order of Mesh.cpp:247-263: bottom face counter-clockwise, then top
    base = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                     [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=float)
    return base * np.array([hx, hy, hz]) + np.array(x0)


def element_vec(u_of_x, X):
    """Element DOF vector in [x0..x7, y0..y7, z0..z7] order."""
    U = np.array([u_of_x(p) for p in X])  # (8,3)
    return np.concatenate([U[:, 0], U[:, 1], U[:, 2]])


def lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


@pytest.mark.parametrize("h", [(1.0, 1.0, 1.0), (0.2, 0.3, 0.5)])
def test_affine_strain_energy_exact(h):
    """u = eps x for a symmetric strain eps: u^T K u = V (lam tr(eps)^2 + 2 mu eps:eps).
    2x2x2 Gauss integration is exact for affine fields on a brick."""
    E, nu = 2.1e5, 0.3
    X = brick(*h, x0=(0.3, -0.1, 0.7))
    K, _ = fe.hex8(X[None], np.array([E]), nu, np.zeros(3))
    K = K[0]
    lam, mu = lame(E, nu)
    V = h[0] * h[1] * h[2]
    rng = np.random.default_rng(1)
    for _ in range(5):
        A = rng.standard_normal((3, 3))
        eps = 0.5 * (A + A.T)
        u = element_vec(lambda p: eps @ p, X)
        energy = lam * np.trace(eps) ** 2 + 2 * mu * np.sum(eps * eps)
        assert u @ K @ u == pytest.approx(V * energy, rel=1e-12)


def test_rigid_modes_in_null_space_and_spectrum():
    X = brick(0.1, 0.1, 0.1, x0=(0.5, 0.5, 0.5))
    K, _ = fe.hex8(X[None], np.array([1000.0]), 0.3, np.zeros(3))
    K = K[0]
    assert np.allclose(K, K.T, rtol=0, atol=1e-12 * np.abs(K).max())
    for t in np.eye(3):
        u = element_vec(lambda p: t, X)
        assert np.linalg.norm(K @ u) <= 1e-12 * np.linalg.norm(K)
    for w in np.eye(3):
        u = element_vec(lambda p: np.cross(w, p), X)
        assert np.linalg.norm(K @ u) <= 1e-12 * np.linalg.norm(K) * np.linalg.norm(u)
    ev = np.linalg.eigvalsh(K)
    assert np.sum(np.abs(ev) < 1e-10 * ev.max()) == 6
    assert ev.min() > -1e-10 * ev.max()


def test_h_scaling_and_load_sum():
    E, nu = 3.0, 0.25
    acc = np.array([0.5, -1.0, 2.0])
    K1, f1 = fe.hex8(brick(1, 1, 1)[None], np.array([E]), nu, acc)
    K2, f2 = fe.hex8(brick(2, 2, 2)[None], np.array([E]), nu, acc)
    assert np.allclose(K2, 2 * K1, rtol=1e-13, atol=0)
    for c in range(3):
        assert f1[0, 8 * c:8 * c + 8].sum() == pytest.approx(acc[c] * 1.0, rel=1e-13)
        assert f2[0, 8 * c:8 * c + 8].sum() == pytest.approx(acc[c] * 8.0, rel=1e-13)


@pytest.mark.parametrize("m_el", [2, 5])
def test_upper_csr_nnz_closed_form(m_el):
    """nnz(upper K_i) = (9(3m-2)^3 + n)/2 with m nodes per axis (27-node stencil)."""
    o = opts(Nx=1, Ny=1, Nz=1, nx=m_el, ny=m_el, nz=m_el)
    s = fe.Subdomains(Mesh(o))
    m = m_el + 1
    n = 3 * m ** 3
    ip, _ = fe.Subdomains.upper_csr_pattern(s.K[0])
    assert len(s.l2g[0]) == n
    assert ip[-1] == (9 * (3 * m - 2) ** 3 + n) // 2


def test_poisson_1d_kat():
    """Matrix::testSolver (Matrix.cpp:1329-1401): tridiag(-1,2,-1), n=5, h=1/6,
    rhs h^2, exact u(x) = x(1-x)/2 at the nodes."""
    from oracle.feti import _Factor
    n, h = 5, 1.0 / 6.0
    A = 2 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    B = np.full((n, 3), h * h)
    x = np.arange(1, n + 1) * h
    for sparse in (False, True):
        U = _Factor(A, sparse).solve(B)
        for c in range(3):
            assert np.allclose(U[:, c], x * (1 - x) / 2, rtol=1e-13)
