"""CPU tests of the native library: exports, and bit-exact integer structures
(numbering, Bf rows/columns/signs, weights, j_col_cmpr, s-sets, Dirichlet and
corner lists) against the oracle's restatement of the reference."""
import ctypes
import os
import re

import numpy as np
import pytest

import hfls_amd
from oracle.mesh import Mesh as OMesh, opts as oopts
from oracle.fe import Subdomains
from oracle.decomp import Gluing
from oracle.kernel import sset_geometric

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    lib = hfls_amd.load()
    hdr = open(os.path.join(ROOT, "include", "hfls_gpu.h")).read()
    names = re.findall(r"^(?:int|const char\*)\s+(hfls_\w+)\(", hdr, flags=re.M)
    assert len(names) >= 20
    assert sorted(names) == sorted(hfls_amd.EXPORTS)
    for n in names:
        assert hasattr(lib, n), n


def test_host_alloc_argument_checks():
    """hfls_host_alloc rejects a size <= 0 and a null output pointer before it asks
    the runtime for memory (no device needed); freeing a null pointer is a no-op."""
    lib = hfls_amd.load()
    p = ctypes.c_void_p(1)
    assert lib.hfls_host_alloc(ctypes.c_int64(0), ctypes.byref(p)) != 0
    assert b"positive" in lib.hfls_last_error()
    assert p.value is None
    assert lib.hfls_host_alloc(ctypes.c_int64(64), None) != 0
    assert b"null" in lib.hfls_last_error()
    assert lib.hfls_host_free(None) == 0


def test_unknown_option_is_an_error():
    with pytest.raises(hfls_amd.HflsError, match="unknown option"):
        hfls_amd.Problem(not_an_option="1")


def test_option_values_of_the_reference():
    """htfeti.cpp:18 lists Dirichlet_explicit, Dirichlet_implicit and lumped; both Dirichlet
    forms build S_bb (Cluster.cpp:140-142).  A value outside the list is an error (the
    reference would silently take it as lumped, Cluster.cpp:1104-1107)."""
    for v in ("Dirichlet_explicit", "Dirichlet_implicit", "lumped"):
        assert hfls_amd.Problem(preconditioner=v, nx=2, ny=2, nz=2).nsub == 8
    with pytest.raises(hfls_amd.HflsError, match="unknown preconditioner"):
        hfls_amd.Problem(preconditioner="Dirichlet")
    with pytest.raises(hfls_amd.HflsError, match="unknown typeBc"):
        hfls_amd.Problem(typeBc="corners")
    # matrix-core format of the mixed-precision block GEMM (no counterpart in the reference)
    for v in ("f16x2", "f32"):
        assert hfls_amd.Problem(block_gemm=v, nx=2, ny=2, nz=2).nsub == 8
    with pytest.raises(hfls_amd.HflsError, match="unknown block_gemm"):
        hfls_amd.Problem(block_gemm="bf16")


CASES = [
    dict(),  # reference defaults: 2x2x2, 5^3, symmetry BC, checkerboard E
    dict(bc="clamp_x0", material="uniform", young_modulus="2.1e5", load="body"),  # 2x2x2 x 5^3 clamped cube, body force
    dict(Nx=3, Ny=2, Nz=1, nx=2, ny=3, nz=2),
    dict(Cx=2, Cy=1, Cz=2, Nx=2, Ny=2, Nz=1, nx=3, ny=2, nz=2, bc="clamp_x0"),
    dict(Cx=1, Cy=2, Cz=1, Nx=2, Ny=1, Nz=2, slabs_x="2,4", slabs_y="3,1", slabs_z="2,3",
         material="lognormal", young_modulus="2.1e5", seed="7", bc="clamp_x0"),
    # uneven batching: per-axis slab sizes drawn from a list with the seed
    dict(Cx=2, Cy=1, Cz=1, Nx=2, Ny=2, Nz=2, slab_choices="2,3,4", material="lognormal", young_modulus="2.1e5",
         seed="5", mat_sigma="1.5", bc="clamp_x0"),
]


@pytest.mark.parametrize("kw", CASES)
def test_integer_structures_bit_exact(kw):
    o = oopts(**kw)
    m = OMesh(o)
    s = Subdomains(m)
    g = Gluing(m, s)
    p = hfls_amd.Problem(**kw)
    assert p.nsub == m.nsub and p.nlam == g.nlam and p.n_glue == g.n_glue and p.nPoints == m.nPoints
    assert np.array_equal(p.dirichlet(), m.dirichlet_dofs)
    assert np.array_equal(p.corners(), m.corner_dofs)
    a, b, dof, w = p.rows()
    assert np.array_equal(a, g.row_a[:, 0]) and np.array_equal(dof, g.row_a[:, 1])
    assert np.array_equal(b[:g.n_glue], g.row_b[:, 0]) and np.all(b[g.n_glue:] == -1)
    assert np.array_equal(w, g.weight)  # same expression 1/((m-1)+1): bitwise equal
    E = m.young_per_sub()
    for d in range(m.nsub):
        assert np.array_equal(p.sub_l2g(d), s.l2g[d])
        rows, dofs, signs = p.sub_bf(d)
        bf = g.Bf[d]
        assert np.array_equal(rows, bf["rows"]) and np.array_equal(dofs, bf["col"])
        assert np.array_equal(signs, bf["val"].astype(int))
        assert np.array_equal(p.sub_bdofs(d), bf["j_col_cmpr"])
        lnodes = s.l2g[d][0::3] // 3
        assert np.array_equal(p.sub_sset(d), sset_geometric(m.grid[lnodes], bf["j_col_cmpr"]))
        info = p.sub_info(d)
        assert info["cluster"] == m.cluster_of(d)
        assert info["n_r"] + info["n_s"] == info["n"] == len(s.l2g[d])
        assert p.sub_young(d) == pytest.approx(E[d], rel=1e-15)


@pytest.mark.slow
def test_config1_integer_hashes():
    """512-subdomain cube decomposition: 8 clusters x 4^3 subdomains of 10^3 elements;
    sizes derived from the reference's construction rules."""
    kw = dict(Cx=2, Cy=2, Cz=2, Nx=4, Ny=4, Nz=4, nx=10, ny=10, nz=10, bc="clamp_x0")
    p = hfls_amd.Problem(**kw)
    assert p.nsub == 512 and p.nPoints == 81 ** 3
    assert p.nlam == 588846
    ns = [p.sub_info(d)["n_s"] for d in range(p.nsub)]
    nlam_i = [p.sub_info(d)["nbf"] for d in range(p.nsub)]
    assert min(nlam_i) == 1191 and max(nlam_i) == 2598
    assert abs(np.mean(nlam_i) - 2254.8) < 0.1
    assert max(ns) == 1806


@pytest.mark.parametrize("kw", [CASES[1], CASES[3], CASES[4]])
def test_traction_loads_bit_exact(kw):
    """The product's seeded face tractions equal the oracle's bit for bit (uneven
    slabs and several clusters included)."""
    from oracle.loads import traction_loads
    o = oopts(**kw)
    m = OMesh(o)
    s = Subdomains(m)
    p = hfls_amd.Problem(**kw)
    F = p.traction(987654321, 2, 3)
    Fo = traction_loads(m, s, 987654321, [2, 3, 4])
    for j in range(3):
        assert np.array_equal(F[j], Fo[j])


def test_slab_draw_uses_only_the_choices_and_all_of_them():
    """The seeded slab draw (uneven batching) picks every listed size and nothing
    else, roughly uniformly, and the product draws exactly the oracle's sizes."""
    from oracle.mesh import _slabs
    o = oopts(Cx=8, Cy=8, Cz=8, Nx=8, Ny=8, Nz=8, slab_choices="6,8,10,12,14", seed="20240611")
    for ax in "xyz":
        v = _slabs(o, ax)
        assert len(v) == 64 and set(v) == {6, 8, 10, 12, 14}
        counts = np.bincount(v)[[6, 8, 10, 12, 14]]
        assert counts.min() >= 4 and counts.max() <= 24
    kw = dict(Cx=2, Cy=2, Cz=2, Nx=4, Ny=4, Nz=4, slab_choices="6,8,10,12,14", seed="20240611")
    m = OMesh(oopts(**kw))
    p = hfls_amd.Problem(**kw)
    assert p.nPoints == m.nPoints
    for d in range(0, m.nsub, 37):
        assert p.sub_info(d)["n"] == 3 * len(np.unique(m.elements[m.part == d]))


def test_mirror_shape_checks():
    """The Python mirror rejects arrays whose size differs from what the C ABI reads."""
    from hfls_amd import _vec, _rows
    assert _vec(np.arange(4.0), 4, "x").flags["C_CONTIGUOUS"]
    assert _rows(np.zeros(5), 5, "F").shape == (1, 5)
    with pytest.raises(ValueError):
        _vec(np.zeros(3), 4, "x")
    with pytest.raises(ValueError):
        _vec(np.zeros((2, 2)), 4, "x")
    with pytest.raises(ValueError):
        _rows(np.zeros((0, 5)), 5, "F")
    with pytest.raises(ValueError):
        _rows(np.zeros((2, 4)), 5, "F")


# Bf glue rows where 2, 4 and 8 subdomains share a DOF, derived by hand from the loop of
# Cluster.cpp:846-903 and the geometry (not from the oracle's grouping): subdomain d visits
# its interface DOFs in ascending global order (skipping Dirichlet DOFs) and, for each,
# the subdomains above d that also hold it in ascending order (the data.interface lists
# filled by the pairwise loop of Cluster.cpp:705-727): one row each, +1 in d, -1 in the
# neighbour, weight 1 / (number of subdomains holding the DOF).  Cube of 2x2x2 subdomains
# of 2x2x2 elements: global node (i, j, k), 0 <= i, j, k <= 4, is i + 5 j + 25 k;
# subdomain (I, J, K) is I + 2 J + 4 K and holds nodes 2I <= i <= 2I + 2 (same in j, k);
# clamp_x0 makes the DOFs of i = 0 Dirichlet.
def _holders(i, j, k):
    side = lambda x: [0] if x < 2 else ([0, 1] if x == 2 else [1])
    return sorted(I + 2 * J + 4 * K for I in side(i) for J in side(j) for K in side(k))


def _hand_rows(d):
    I, J, K = d & 1, (d >> 1) & 1, d >> 2
    nodes = sorted((i + 5 * j + 25 * k, (i, j, k)) for i in range(2 * I, 2 * I + 3) for j in range(2 * J, 2 * J + 3)
                   for k in range(2 * K, 2 * K + 3))
    rows = []
    for node, (i, j, k) in nodes:
        h = _holders(i, j, k)
        if len(h) == 1 or i == 0:  # not on an interface, or Dirichlet
            continue
        for c in range(3):
            rows += [(d, s, 3 * node + c, 1.0 / len(h)) for s in h if s > d]
    return rows


def test_bf_rows_of_dofs_shared_by_4_and_8_subdomains():
    """Row order, signs and weights of the Bf glue rows of subdomains 0 and 1 (where DOFs
    are shared by 2, 4 and 8 subdomains) against the hand derivation above, for the
    native library and the oracle; spot check of the rows of the centre node (all 8
    subdomains hold it: 7 rows from subdomain 0, 6 from subdomain 1, weight 1/8)."""
    kw = dict(Nx=2, Ny=2, Nz=2, nx=2, ny=2, nz=2, bc="clamp_x0")
    hand = _hand_rows(0) + _hand_rows(1)
    n0 = len(_hand_rows(0))
    centre = 3 * (2 + 5 * 2 + 25 * 2)
    assert [r for r in hand if r[2] == centre] == [(0, s, centre, 0.125) for s in range(1, 8)] + \
        [(1, s, centre, 0.125) for s in range(2, 8)]
    p = hfls_amd.Problem(**kw)
    a, b, dof, w = p.rows()
    got = list(zip(a[:len(hand)].tolist(), b[:len(hand)].tolist(), dof[:len(hand)].tolist(), w[:len(hand)].tolist()))
    assert got[:n0] == hand[:n0] and got == hand
    assert a[len(hand)] == 2  # the rows of subdomain 2 follow
    m = OMesh(oopts(**kw))
    g = Gluing(m, Subdomains(m))
    ogot = list(zip(g.row_a[:len(hand), 0].tolist(), g.row_b[:len(hand), 0].tolist(),
                    g.row_a[:len(hand), 1].tolist(), g.weight[:len(hand)].tolist()))
    assert ogot == hand
    # every pair of the 8 holders of the centre DOF glued once: 28 rows
    assert int(np.sum(dof[:p.n_glue] == centre)) == 28
