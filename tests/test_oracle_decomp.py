"""Pins of the oracle's mesh numbering and gluing structures (integer work)."""
import math

import numpy as np
import pytest

from oracle.mesh import Mesh, opts
from oracle.fe import Subdomains
from oracle.decomp import Gluing, build_Bc, bc_rank_full
from oracle.feti import HTFETI


def test_mesh_counts_and_partition():
    o = opts(Nx=2, Ny=3, Nz=2, nx=2, ny=1, nz=3)
    m = Mesh(o)
    NX, NY, NZ = 4, 3, 6
    assert m.nPoints == (NX + 1) * (NY + 1) * (NZ + 1)
    assert len(m.elements) == NX * NY * NZ
    # each subdomain owns nx*ny*nz elements; ids II + JJ*Nx + KK*Nx*Ny
    assert np.all(np.bincount(m.part) == 2 * 1 * 3)
    e = 3 + 2 * NX + 5 * NX * NY  # element (ii=3, jj=2, kk=5)
    assert m.part[e] == 1 + 2 * 2 + 1 * 6
    # corner nodes: subdomain-grid vertices minus the 8 cube corners
    assert len(m.corner_dofs) == 3 * ((2 + 1) * (3 + 1) * (2 + 1) - 8)
    # remainder(p, 2) pattern 0, 1, 0, -1
    assert list(m.material_id[:8]) == [0, 1, 0, -1, 0, 1, 0, -1]


def test_dirichlet_lists():
    m = Mesh(opts(Nx=1, Ny=1, Nz=1, nx=2, ny=2, nz=2, bc="symmetry"))
    # 9 nodes on each coordinate plane, one component each
    assert len(m.dirichlet_dofs) == 27
    assert list(m.dirichlet_dofs[:4]) == [0, 1, 2, 3 * 1 + 1]
    assert np.all(np.diff(m.dirichlet_dofs) > 0)
    m2 = Mesh(opts(Nx=1, Ny=1, Nz=1, nx=2, ny=2, nz=2, bc="clamp_x0"))
    assert len(m2.dirichlet_dofs) == 27
    assert set(m2.dirichlet_dofs // 3) == {i for i in range(27) if i % 3 == 0}


@pytest.mark.parametrize("kw", [dict(), dict(bc="clamp_x0"), dict(Nx=3, Ny=2, Nz=1, nx=2, ny=3, nz=2),
                                dict(Cx=2, Nx=2, Ny=1, Nz=2, nx=2, ny=2, nz=2)])
def test_bf_closed_forms(kw):
    o = opts(**kw)
    m = Mesh(o)
    s = Subdomains(m)
    g = Gluing(m, s)
    mult = g.mult
    nondir = np.ones(len(mult), dtype=bool)
    nondir[m.dirichlet_dofs] = False
    n_glue = int(np.sum((mult * (mult - 1) // 2)[nondir]))
    n_dir = sum(int(np.isin(m.dirichlet_dofs, s.l2g[d]).sum()) for d in range(m.nsub))
    assert g.nlam == n_glue + n_dir
    # every glue row: +1 on the lower subdomain, -1 on the higher, same global DOF
    rows = np.zeros(g.nlam, dtype=int)
    for d in range(m.nsub):
        bf = g.Bf[d]
        gl = s.l2g[d][bf["col"]]
        assert np.all(np.diff(bf["rows"]) > 0)
        rows[bf["rows"]] += 1
        for r, gg, v in zip(bf["rows"], gl, bf["val"]):
            if r < g.n_glue:
                a = g.row_a[r]
                b = g.row_b[r]
                assert gg == a[1] == b[1]
                assert (d == a[0] and v == 1) or (d == b[0] and v == -1)
                assert a[0] < b[0]
            else:
                assert v == 1 and gg in set(m.dirichlet_dofs)
    assert np.all(rows[:g.n_glue] == 2) and np.all(rows[g.n_glue:] == 1)
    # weights 1/multiplicity
    assert np.allclose(g.weight[:g.n_glue], 1.0 / mult[g.row_a[:g.n_glue, 1]])
    # rows are ordered by (d, g, s)
    key = g.row_a[:g.n_glue, 0] * 10 ** 12 + g.row_a[:g.n_glue, 1] * 10 ** 4 + g.row_b[:, 0]
    assert np.all(np.diff(key) > 0)


def test_two_subdomain_hand_case():
    """Nx=2, one element each: the shared face x=1/2 has 4 nodes = 12 DOFs, each
    glued once; clamp_x0 adds the 12 DOFs of x=0 in subdomain 0."""
    o = opts(Nx=2, Ny=1, Nz=1, nx=1, ny=1, nz=1, bc="clamp_x0")
    m = Mesh(o)
    s = Subdomains(m)
    g = Gluing(m, s)
    assert g.n_glue == 12 and g.nlam == 24
    shared = [3 * n + c for n in (1, 4, 7, 10) for c in range(3)]
    assert list(g.row_a[:12, 1]) == shared
    assert list(g.row_a[12:, 1]) == [3 * n + c for n in (0, 3, 6, 9) for c in range(3)]
    assert np.all(g.row_a[12:, 0] == 0)


def test_bc_ker_keeps_faces_only():
    o = opts(eps_iter="1e-8")
    h = HTFETI(o)
    # 2x2x2 subdomains: 12 face pairs (rank 6) kept, edge and corner pairs rejected
    assert h.clusters[0]["nlc"] == 72
    A = np.ones((6, 3))
    assert not bc_rank_full(A)
    assert bc_rank_full(np.linalg.qr(np.random.default_rng(0).standard_normal((30, 6)))[0].T)


def test_bc_cor_counts():
    o = opts(typeBc="cor")
    m = Mesh(o)
    s = Subdomains(m)
    g = Gluing(m, s)
    R = [np.eye(len(s.l2g[d]), 6) for d in range(m.nsub)]
    bc = build_Bc(m, s, g, R, "cor", True)
    # one row per (d, corner DOF) with a higher neighbour (full_rank break):
    # multiplicity - 1 rows per corner DOF; 2x2x2: centre 7, face centres 3,
    # edge midpoints 1 (x3 components) = 21 + 54 + 36
    n = sum(int(g.mult[gg]) - 1 for gg in m.corner_dofs if g.mult[gg] > 1)
    assert bc[0]["nlam"] == n == 111
