"""Pins of the oracle's kernel detection, null pivots and regularisation."""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle.mesh import Mesh, opts
from oracle.fe import Subdomains
from oracle.decomp import Gluing
from oracle.kernel import (kernel_from_K, null_pivots, projector_pivots, sset_geometric,
                           regularize, gram_schmidt)


def analytic_rbm(points):
    """Six rigid-body modes (Data.cpp:61-109 construction, own code)."""
    n = len(points)
    R = np.zeros((3 * n, 6))
    x, y, z = points[:, 0], points[:, 1], points[:, 2]
    R[0::3, 0] = 1
    R[1::3, 1] = 1
    R[2::3, 2] = 1
    R[0::3, 3], R[1::3, 3] = -y, x
    R[0::3, 4], R[2::3, 4] = -z, x
    R[1::3, 5], R[2::3, 5] = -z, y
    return R


@pytest.fixture(scope="module")
def small():
    o = opts(Nx=2, Ny=2, Nz=1, nx=3, ny=2, nz=3, bc="clamp_x0")
    m = Mesh(o)
    s = Subdomains(m)
    g = Gluing(m, s)
    return m, s, g


@pytest.mark.parametrize("mode", ["reference", "geometric"])
def test_kernel_detection(small, mode):
    m, s, g = small
    for d in range(m.nsub):
        K = s.K[d].toarray()
        if mode == "reference":
            R, dfc, _ = kernel_from_K(K, None, seed=d)
        else:
            lnodes = s.l2g[d][0::3] // 3
            ss = sset_geometric(m.grid[lnodes], g.Bf[d]["j_col_cmpr"])
            R, dfc, _ = kernel_from_K(K, ss)
        assert dfc == 6
        assert np.linalg.norm(K @ R) <= 1e-10 * np.linalg.norm(K) * np.linalg.norm(R)
        assert np.allclose(R.T @ R, np.eye(6), atol=1e-12)
        Ra = analytic_rbm(m.points[s.l2g[d][0::3] // 3])
        Qa = np.linalg.qr(Ra)[0]
        # sine of the largest principal angle between span(R) and span(RBM)
        sin_max = np.linalg.norm(R - Qa @ (Qa.T @ R), 2)
        assert sin_max <= 1e-10


def test_regularisation_is_generalised_inverse(small):
    m, s, g = small
    for d in range(m.nsub):
        K = s.K[d].toarray()
        lnodes = s.l2g[d][0::3] // 3
        ss = sset_geometric(m.grid[lnodes], g.Bf[d]["j_col_cmpr"])
        R, _, _ = kernel_from_K(K, ss)
        for fix in (null_pivots(R), ss[projector_pivots(R[ss])]):
            assert len(fix) == 6 and abs(np.linalg.det(R[fix])) > 1e-12
            Kr = regularize(K, fix)
            KpK = sla.cho_solve(sla.cho_factor(Kr), K)
            assert np.linalg.norm(K - K @ KpK) <= 1e-10 * np.linalg.norm(K)


def test_null_pivots_hand_case():
    # identity rows 2, 5 and a dominant entry in row 4 (column 1)
    R = np.zeros((6, 2))
    R[2, 0] = 1.0
    R[4, 1] = 3.0
    R[5, 0] = 0.5
    assert list(null_pivots(R)) == [2, 4]
    # the last maximum in column-major order wins on ties (compareDouble is <=)
    T = np.array([[1.0], [1.0], [-1.0]])
    assert list(null_pivots(T)) == [2]


def test_projector_pivots_basis_independent():
    rng = np.random.default_rng(3)
    Rs = np.linalg.qr(rng.standard_normal((24, 6)))[0]
    Q = np.linalg.qr(rng.standard_normal((6, 6)))[0]
    assert list(projector_pivots(Rs)) == list(projector_pivots(Rs @ Q))


def test_gram_schmidt():
    A = np.random.default_rng(5).standard_normal((40, 6))
    Q = gram_schmidt(A)
    assert np.allclose(Q.T @ Q, np.eye(6), atol=1e-13)
    assert np.allclose(np.triu(Q.T @ A), Q.T @ A, atol=1e-12)
