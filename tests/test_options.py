"""The reference's options map at the C ABI (no GPU: host problem only), NULL
handles, and the oracle pins of the option-selected variants.

* every key of htfeti.cpp:12-57 is accepted with the reference's values; the
  values a path cannot honour are rejected with the reason (METIS partitions,
  Dissection, per-subdomain dumps); Ac_clust_explicit=false (the Sc branch) is
  accepted as the two-stage coarse solve;
* metis=true with nparts=1: one subdomain over the whole cube (Mesh.cpp:413-421),
  integer structures bit-exact vs the oracle, material ids of the geometric
  partition in the VTK output;
* oracle pins: corner plus averaged-edge gluing has dim ker(Gc^T Gc) = 6 and the
  displacement of `ker`; the regularised (non-extended) Ac and the two-stage
  Schur-complement solve of Ac_clust_explicit=false give the same displacement as
  the extended Ac; the two-stage solve equals the inverse of [Fc Gc; Gc^T -H H^T]
  (block-inverse identity, the form the device builds); the single-subdomain map
  equals the global direct solve;
* a NULL handle or output pointer is an error return, not a crash.
"""
import ctypes
import re

import numpy as np
import pytest

import hfls_amd
from oracle.mesh import opts, Mesh as OMesh
from oracle.fe import Subdomains
from oracle.decomp import Gluing
from oracle.feti import HTFETI
from oracle.direct import global_solve

REFERENCE_MAP = dict(eps_iter="1e-6", max_iter="200", vtkWithinIter="true", metis="true",
                     preconditioner="Dirichlet_implicit", Ac_clust_explicit="true", path2data="data/",
                     young_modulus="1000", ratio_mat="1e4", poissons_ratio="0.3", linear_solver="pardiso",
                     print_matrices="1", typeBc="ker", Ac_extended_by_kerGc="false",
                     GcTGc_assembl_block_by_block="true", Bc_fullRank="true", create_analytic_ker_K="",
                     Nx="2", Ny="1", Nz="1", nx="5", ny="5", nz="5", nparts="1")
SMALL = dict(Nx=2, Ny=2, Nz=2, nx=3, ny=3, nz=3, bc="clamp_x0", material="uniform", young_modulus="2.1e5",
             load="body")


def test_reference_map_accepted():
    p = hfls_amd.Problem(**REFERENCE_MAP)
    assert p.nsub == 1 and p.ncl == 1
    assert p.nPoints == 11 * 6 * 6


@pytest.mark.parametrize("key,value,match", [
    ("linear_solver", "dissection", "dissection"),
    ("linear_solver", "mumps", "unknown linear_solver"),
    ("nparts", "4", "METIS"),
    ("Ac_clust_explicit", "0", "Ac_clust_explicit"),
    ("print_matrices", "2", "per-subdomain"),
    ("typeBc", "edges", "unknown typeBc"),
    ("vtkWithinIter", "yes", "vtkWithinIter"),
    ("Ac_extended_by_kerGc", "1", "Ac_extended_by_kerGc"),
])
def test_rejected_values(key, value, match):
    with pytest.raises(hfls_amd.HflsError, match=match):
        hfls_amd.Problem(**dict(REFERENCE_MAP, **{key: value}))


def test_metis_needs_uniform_grid_and_one_cluster():
    with pytest.raises(hfls_amd.HflsError, match="one cluster"):
        hfls_amd.Problem(**dict(REFERENCE_MAP, Cx="2"))
    with pytest.raises(hfls_amd.HflsError, match="uniform"):
        hfls_amd.Problem(**dict(REFERENCE_MAP, slabs_x="4,6"))


def test_metis_single_subdomain_matches_oracle():
    import sys
    import os
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    from golden import native_arrays, oracle_arrays
    kw = dict(REFERENCE_MAP)
    a, b = native_arrays(kw), oracle_arrays(kw)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    # corner DOFs come from the geometric 2x1x1 grid that createMesh built
    assert len(a["corners"]) == 3 * 4


def test_metis_vtk_material_ids(tmp_path):
    p = hfls_amd.Problem(**REFERENCE_MAP)
    p.write_vtk(np.zeros(3 * p.nPoints), tmp_path)
    txt = (tmp_path / "box.vtk").read_text()
    dec = txt.split("LOOKUP_TABLE decomposition\n")[1].split("SCALARS")[0].split()
    mat = txt.split("LOOKUP_TABLE materialId\n")[1].split()
    m = OMesh(opts(**REFERENCE_MAP))
    assert all(int(x) == 0 for x in dec)
    assert [int(x) for x in mat] == m.geo_material_id[m.geo_part].tolist()
    assert set(int(x) for x in mat) == {0, 1}


def test_cor_edge_counts_match_oracle():
    """The product's corner plus averaged-edge rows are built on the device side of
    the set-up; here the oracle's rule is pinned by its invariants."""
    h = HTFETI(opts(typeBc="cor+edge", **SMALL))
    for c in h.clusters:
        ev = np.linalg.eigvalsh(c["Gc"].T @ c["Gc"])
        assert np.sum(ev < 1e-10 * ev.max()) == 6
    hk = HTFETI(opts(**SMALL))
    u_e = h.pcpg(eps=1e-11, max_iter=300)["u"]
    u_k = hk.pcpg(eps=1e-11, max_iter=300)["u"]
    assert np.linalg.norm(u_e - u_k) <= 1e-9 * np.linalg.norm(u_k)
    # row count: corners as `cor` plus 3 rows per (edge, first neighbour): a 2x2x2 cluster of
    # 3^3-element subdomains has 6 shared vertex lines... count by hand from the rule
    m = OMesh(opts(**SMALL))
    s = Subdomains(m)
    g = Gluing(m, s)
    from oracle.decomp import build_Bc
    R = [np.zeros((len(l), 6)) for l in s.l2g]
    cor = build_Bc(m, s, g, R, "cor", True)[0]["nlam"]
    ce = build_Bc(m, s, g, R, "cor+edge", True)[0]["nlam"]
    # the 2x2x2 grid has 12 interior half-edges on each of the 3 axes inside the cube
    # shared by 2 or 4 subdomains: edges lying on the 3 internal planes; every edge
    # shared by s subdomains gives s - 1 chained pairs, 3 components each
    edges = 0
    for a in range(3):
        # edges along axis a: positions of the other two coordinates on the 3x3 vertex
        # lines of the subdomain grid, 2 segments each along a
        for pb in range(3):
            for pc in range(3):
                sharing = (2 if pb == 1 else 1) * (2 if pc == 1 else 1)
                edges += 2 * (sharing - 1)
    assert ce - cor == 3 * edges


def test_regularised_ac_same_displacement():
    kw = dict(SMALL, Cx=2, Nx=2)
    ue = HTFETI(opts(**kw)).pcpg(eps=1e-11, max_iter=300)["u"]
    ur = HTFETI(opts(Ac_extended_by_kerGc="false", **kw)).pcpg(eps=1e-11, max_iter=300)["u"]
    assert np.linalg.norm(ue - ur) <= 1e-10 * np.linalg.norm(ue)


def test_two_stage_coarse_solve():
    """Ac_clust_explicit=false: beta = (Sc + H H^T)^-1 (Gc^T Fc^-1 d - e), mu = Fc^-1 (d -
    Gc beta) with Sc = Gc^T Fc^-1 Gc (myModul.py:128-135) is the inverse of
    [Fc Gc; Gc^T -H H^T], and the solve gives the displacement of the extended Ac."""
    kw = dict(SMALL, Cx=2, Nx=2)
    h = HTFETI(opts(Ac_clust_explicit="false", **kw))
    p = hfls_amd.Problem(**dict(kw, Ac_clust_explicit="false"))
    assert p.nsub == h.nsub
    rng = np.random.default_rng(11)
    for c in h.clusters:
        nlc, nr, Fc, Gc, H = c["nlc"], c["nr"], c["Fc"], c["Gc"], c["H"]
        iF0, iS0, Sc = c["two_stage"]
        assert np.allclose(Sc, Gc.T @ np.linalg.solve(Fc, Gc), rtol=1e-12, atol=1e-12 * np.abs(Sc).max())
        A = np.block([[Fc, Gc], [Gc.T, -H @ H.T]])
        b = rng.standard_normal(nlc + nr)
        import scipy.linalg as sla
        beta = sla.cho_solve(iS0, Gc.T @ sla.cho_solve(iF0, b[:nlc]) - b[nlc:])
        mu = sla.cho_solve(iF0, b[:nlc] - Gc @ beta)
        ref = np.linalg.solve(A, b)
        assert np.linalg.norm(np.concatenate([mu, beta]) - ref) <= 1e-10 * np.linalg.norm(ref)
    ue = HTFETI(opts(**kw)).pcpg(eps=1e-11, max_iter=300)["u"]
    us = h.pcpg(eps=1e-11, max_iter=300)["u"]
    assert np.linalg.norm(ue - us) <= 1e-10 * np.linalg.norm(ue)


def test_metis_single_subdomain_oracle_vs_direct():
    h = HTFETI(opts(**REFERENCE_MAP))
    assert h.nsub == 1
    u = h.pcpg(eps=1e-12, max_iter=200)["u"]
    ud, _, _ = global_solve(h.mesh)
    assert np.linalg.norm(u - ud) <= 1e-9 * np.linalg.norm(ud)
    # the geometric 2x1x1 decomposition of the same cube and material gives the same u
    ug = HTFETI(opts(**dict(REFERENCE_MAP, metis="false"))).pcpg(eps=1e-12, max_iter=200)["u"]
    assert np.linalg.norm(u - ug) <= 1e-9 * np.linalg.norm(ud)


def test_null_handles_are_errors():
    lib = hfls_amd.load()
    buf = np.zeros(16, dtype=np.int64)
    ptr = buf.ctypes.data_as(ctypes.c_void_p)
    assert lib.hfls_problem_sizes(None, ptr) != 0
    assert b"null problem handle" in lib.hfls_last_error()
    assert lib.hfls_mesh_sizes(None, ptr) != 0
    assert b"null mesh handle" in lib.hfls_last_error()
    assert lib.hfls_apply_F(None, ptr, ptr) != 0
    assert lib.hfls_server_step(None, None) != 0
    assert b"null server handle" in lib.hfls_last_error()
    assert lib.hfls_server_cancel(None, ctypes.c_int64(1)) != 0
    p = hfls_amd.Problem(**SMALL)
    assert lib.hfls_problem_sizes(p._h, None) != 0
    assert b"null argument" in lib.hfls_last_error()
    assert lib.hfls_problem_create(0, None, None, None) != 0
    # destroying NULL handles is a no-op
    assert lib.hfls_problem_destroy(None) == 0 and lib.hfls_mesh_destroy(None) == 0
