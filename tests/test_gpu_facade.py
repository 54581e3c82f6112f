"""The drop-in boundary on the GPU: the reference's own options map through the C++
facade, and the Cluster operators the facade forwards (Cluster.hpp:20-121).

* hfls_amd/htfeti --reference-map runs main() of htfeti.cpp with its options map
  value for value (metis=true with nparts=1, Dirichlet_implicit, pardiso,
  print_matrices=1, Ac_extended_by_kerGc=false, vtkWithinIter=true, path2data):
  it converges, writes box<it>.vtk after every iteration and box.vtk at the end,
  and the MatrixMarket dumps of the cluster matrices; the displacements in box.vtk
  equal the oracle's for the same map to the file's 6 decimals;
* mult_Gf, mult_GfT, mult_RfT and Rf against each other and the oracle's span
  of Rf (they depend on the basis of R and H, so the checks are basis free:
  Gf a = Bf (Rf a), adjointness, span), the recovery of printVTK against the
  solve (1e-9) and the oracle (1e-8), and the getters of Cluster.hpp:26-31.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import hfls_amd
from oracle.mesh import opts
from oracle.feti import HTFETI

pytestmark = pytest.mark.gpu

# htfeti.cpp:12-57 as the library sees it (path2data aside)
REFERENCE_MAP = dict(eps_iter="1e-6", max_iter="200", vtkWithinIter="true", metis="true",
                     preconditioner="Dirichlet_implicit", Ac_clust_explicit="true", young_modulus="1000",
                     ratio_mat="1e4", poissons_ratio="0.3", linear_solver="pardiso", print_matrices="1",
                     typeBc="ker", Ac_extended_by_kerGc="false", GcTGc_assembl_block_by_block="true",
                     Bc_fullRank="true", Nx="2", Ny="1", Nz="1", nx="5", ny="5", nz="5", nparts="1")
MULTI = dict(Cx=2, Cy=1, Cz=1, Nx=2, Ny=2, Nz=2, nx=3, ny=3, nz=4, bc="clamp_x0", material="lognormal",
             young_modulus="2.1e5", seed="3", mat_sigma="0.5", load="body")


def read_vtk_displacements(path):
    with open(path) as f:
        txt = f.read()
    m = re.search(r"POINT_DATA (\d+)\nSCALARS displacements float 3\nLOOKUP_TABLE my_table\n", txt)
    n = int(m.group(1))
    vals = txt[m.end():].split("\n")[:n]
    return np.array([[float(x) for x in v.split()] for v in vals]).ravel()


def read_mm(path):
    with open(path) as f:
        head = f.readline()
        rows, cols, nnz = (int(x) for x in f.readline().split())
        A = np.zeros((rows, cols))
        for _ in range(nnz):
            i, j, v = f.readline().split()
            A[int(i) - 1, int(j) - 1] = float(v)
    if "symmetric" in head:
        A = A + np.triu(A, 1).T
    return A


def test_reference_map_through_the_facade(tmp_path):
    exe = os.path.join(os.path.dirname(hfls_amd.LIB_PATH), "htfeti")
    print("htfeti --reference-map", tmp_path, flush=True)
    r = subprocess.run([exe, "--reference-map", str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and "converged 1" in r.stdout
    its = int(re.search(r"iterations (\d+)", r.stdout).group(1))
    assert "subdomains 1 " in r.stdout
    files = sorted(os.listdir(tmp_path))
    print(files, flush=True)
    for it in range(its):
        assert "box%d.vtk" % it in files
    assert "box%d.vtk" % its not in files
    for name in ("Fc_clust_0", "Gc_clust_0", "GcTGc_clust_0", "kerGc_0", "Ac_clust_0", "weigth"):
        assert "dump_%s.txt" % name in files
    # one subdomain: no cluster constraints, Gc^T Gc = 0 and kerGc = I (6)
    assert np.allclose(read_mm(tmp_path / "dump_kerGc_0.txt"), np.eye(6))
    assert read_mm(tmp_path / "dump_Ac_clust_0.txt").shape == (6, 6)
    ora = HTFETI(opts(**REFERENCE_MAP))
    ro = ora.pcpg(eps=1e-12, max_iter=200)
    u = read_vtk_displacements(tmp_path / "box.vtk")
    assert u.shape == ro["u"].shape
    err = np.abs(u - ro["u"]).max()
    print("box.vtk vs oracle: max abs %.2e (|u|max %.3e, %%f format)" % (err, np.abs(ro["u"]).max()), flush=True)
    assert err <= 1.5e-6
    w = np.loadtxt(tmp_path / "dump_weigth.txt")
    assert np.array_equal(w, ora.glue.weight)


def test_reference_map_mesh_matches_oracle(tmp_path):
    """The same map through the Python mirror (path2data redirected), solved to 1e-10."""
    kw = dict(REFERENCE_MAP, path2data=str(tmp_path), vtkWithinIter="false", print_matrices="0")
    gpu = hfls_amd.Mesh(**kw)
    try:
        r = gpu.pcpg(eps=1e-10, max_iter=200)
        ora = HTFETI(opts(**REFERENCE_MAP))
        ro = ora.pcpg(eps=1e-10, max_iter=200)
        err = np.linalg.norm(r["u"] - ro["u"]) / np.linalg.norm(ro["u"])
        print("single subdomain: iterations", r["iterations"], ro["iterations"], "err %.2e" % err, flush=True)
        assert r["converged"] and abs(r["iterations"] - ro["iterations"]) <= 2 and err <= 1e-8
        assert os.path.exists(tmp_path / "box.vtk") and not os.path.exists(tmp_path / "box0.vtk")
    finally:
        gpu.close()


def test_coarse_operators_and_recovery():
    gpu = hfls_amd.Mesh(**MULTI)
    try:
        ora = HTFETI(opts(**MULTI))
        c = gpu.counts()
        print("counts", c, flush=True)
        assert c["nSubClst"] == ora.nsub and c["nLam_f"] == ora.nlam and c["nRBM_f"] == ora.ncol_G
        assert c["nLam_c"] == sum(cl["nlc"] for cl in ora.clusters) and c["nRBM_c"] == 6 * ora.nsub
        assert c["neqClst"] == sum(len(l) for l in ora.subs.l2g)
        rng = np.random.default_rng(3)
        # Gf and Rf are basis dependent (R, H); compare what does not depend on the basis:
        # the projector they define and Gf = Bf Rf
        Rf = gpu.Rf()
        lam = rng.standard_normal(ora.nlam)
        a = rng.standard_normal(ora.ncol_G)
        Ga = gpu.mult_Gf(a)
        GTl = gpu.mult_GfT(lam)
        assert abs(lam @ Ga - a @ GTl) <= 1e-12 * np.linalg.norm(lam) * np.linalg.norm(Ga)
        # Gf a = Bf (Rf a) per subdomain
        xs, off = [], 0
        for d, l2g in enumerate(ora.subs.l2g):
            n = len(l2g)
            Rd = Rf[6 * off:6 * (off + n)].reshape(6, n).T
            cl = ora.mesh.cluster_of(d)
            xs.append(Rd @ a[6 * cl:6 * cl + 6])
            off += n
        assert np.linalg.norm(ora.mult_Bf(xs) - Ga) <= 1e-12 * np.linalg.norm(Ga)
        # span: Rf_d spans the oracle's Rf_d columns of the same cluster
        off = 0
        for d, l2g in enumerate(ora.subs.l2g):
            n = len(l2g)
            Rd = Rf[6 * off:6 * (off + n)].reshape(6, n).T
            cl = ora.mesh.cluster_of(d)
            Ro = ora.Rf[d][:, 6 * cl:6 * cl + 6]
            Q, _ = np.linalg.qr(Ro)
            assert np.linalg.norm(Rd - Q @ (Q.T @ Rd)) <= 1e-9 * max(1.0, np.linalg.norm(Rd))
            off += n
        # Rf^T f: e for a load f, compared through G (G e' computed both ways)
        f = np.concatenate([rng.standard_normal(len(l)) for l in ora.subs.l2g])
        e = gpu.mult_RfT(f)
        e_chk = np.zeros(ora.ncol_G)
        off = 0
        for d, l2g in enumerate(ora.subs.l2g):
            n = len(l2g)
            Rd = Rf[6 * off:6 * (off + n)].reshape(6, n).T
            cl = ora.mesh.cluster_of(d)
            e_chk[6 * cl:6 * cl + 6] += Rd.T @ f[off:off + n]
            off += n
        assert np.linalg.norm(e - e_chk) <= 1e-12 * np.linalg.norm(e_chk)
        # recovery of printVTK from lambda and the alpha of Projection equals the solve's u
        r = gpu.pcpg(eps=1e-10, max_iter=300)
        fb = gpu.load_vector()
        # alpha of the final gradient: g = F lambda - d with d = Bf K_c^+ f
        fs, off = [], 0
        for l2g in ora.subs.l2g:
            fs.append(fb[off:off + len(l2g)])
            off += len(l2g)
        kf, off = gpu.mult_Kplus_f(fb), 0
        xs = []
        for l2g in ora.subs.l2g:
            xs.append(kf[off:off + len(l2g)])
            off += len(l2g)
        d_rhs = ora.mult_Bf(xs)  # the signed +-1 sums of Bf (integer structure)
        _, alpha = gpu.projection(gpu.mult_Ff(r["lambda"]) - d_rhs)
        x, u = gpu.recover(fb, r["lambda"], alpha)
        err = np.linalg.norm(u - r["u"]) / np.linalg.norm(r["u"])
        print("recovery vs solve %.2e" % err, flush=True)
        assert err <= 1e-9
        uo = ora.pcpg(f=fs, eps=1e-10, max_iter=300)["u"]
        assert np.linalg.norm(u - uo) <= 1e-8 * np.linalg.norm(uo)
        # the per-subdomain pieces scatter to u (last subdomain holding a DOF writes it)
        ug = np.zeros_like(u)
        off = 0
        for l2g in ora.subs.l2g:
            ug[l2g] = x[off:off + len(l2g)]
            off += len(l2g)
        assert np.array_equal(ug, u)
    finally:
        gpu.close()


def test_vtk_within_iterations(tmp_path):
    kw = dict(MULTI, path2data=str(tmp_path), vtkWithinIter="true")
    gpu = hfls_amd.Mesh(**kw)
    try:
        r = gpu.pcpg(eps=1e-8, max_iter=300)
        files = os.listdir(tmp_path)
        print("iterations", r["iterations"], "files", len(files), flush=True)
        assert all("box%d.vtk" % i in files for i in range(r["iterations"]))
        assert "box.vtk" in files and len(files) == r["iterations"] + 1
        # the per-iteration path computes the same solve as the graph-replayed one
        gpu2 = hfls_amd.Mesh(**MULTI)
        r2 = gpu2.pcpg(eps=1e-8, max_iter=300)
        gpu2.close()
        assert r2["iterations"] == r["iterations"] and np.array_equal(r2["u"], r["u"])
        last = read_vtk_displacements(tmp_path / ("box%d.vtk" % (r["iterations"] - 1)))
        assert np.abs(last - read_vtk_displacements(tmp_path / "box.vtk")).max() == 0.0
    finally:
        gpu.close()
