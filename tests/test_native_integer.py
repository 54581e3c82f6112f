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
