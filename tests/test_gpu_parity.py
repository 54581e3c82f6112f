"""GPU parity of the HIP path (through the C ABI) against the CPU oracle.

Tolerances (FP64 everywhere on this path):
  * integer results (defect, fixing DOFs): exact
  * K_i R_i ~ 0 and span(R) vs the oracle's kernel: 1e-9
  * load vector: 1e-13 relative
  * one application of F, P, M, K_c^+: 1e-9 relative (different but equally valid
    factorisation orders of the same K_reg)
  * PCPG: iterations within +-2 of the oracle at the same eps, displacement within
    1e-8 relative of the oracle and of the global direct solve
"""
import numpy as np
import pytest

import hfls_amd
from oracle.mesh import opts
from oracle.feti import HTFETI
from oracle.direct import global_solve

pytestmark = pytest.mark.gpu

CUBE8_5E3 = dict(Nx=2, Ny=2, Nz=2, nx=5, ny=5, nz=5, bc="clamp_x0", material="uniform",
            young_modulus="2.1e5", load="body", body_force="-1.0")
MULTI = dict(Cx=2, Cy=1, Cz=1, Nx=2, Ny=2, Nz=2, nx=3, ny=3, nz=4, bc="clamp_x0", material="lognormal",
             young_modulus="2.1e5", seed="3", mat_sigma="0.5", load="body")
# uneven batching in small: slab sizes drawn from {2,3,4}, lognormal E (sigma 1.5, contrast 1e3)
UNEVEN = dict(Cx=2, Cy=1, Cz=1, Nx=2, Ny=2, Nz=2, slab_choices="2,3,4", bc="clamp_x0", material="lognormal",
              young_modulus="2.1e5", seed="5", mat_sigma="1.5", mat_contrast="1e3", load="body")
CASES = {"cube8_5e3": CUBE8_5E3, "multi": MULTI, "refdefault": dict(nx=3, ny=3, nz=3, ratio_mat="10"), "uneven": UNEVEN,
         # the other gluing and preconditioner of the reference (Cluster.cpp:846-903, 1106-1107)
         "cor": dict(MULTI, typeBc="cor"), "lumped": dict(CUBE8_5E3, preconditioner="lumped"),
         # the 2x2x2 x 5^3 cube with corner plus averaged-edge gluing, and the same gluing on two clusters
         "cube8_5e3_cor_edge": dict(CUBE8_5E3, typeBc="cor+edge"), "multi_cor_edge": dict(MULTI, typeBc="cor+edge"),
         # the reference's default cluster inverse: the singular Ac regularised by its kernel
         "ac_regularized": dict(MULTI, Ac_extended_by_kerGc="false"),
         # Ac_clust_explicit=false: the Sc branch completed as the two-stage coarse solve
         "ac_two_stage": dict(MULTI, Ac_clust_explicit="false"),
         # the reference's own map: metis=true with nparts=1 makes the 10x5x5 cube one
         # subdomain (Mesh.cpp:413-421), checkerboard material of the geometric 2x1x1
         # partition, symmetry planes, the reference body force
         "metis_single": dict(Nx=2, Ny=1, Nz=1, nx=5, ny=5, nz=5, metis="true", nparts="1")}


@pytest.fixture(scope="module", params=sorted(CASES))
def pair(request):
    kw = CASES[request.param]
    print("case", request.param, kw, flush=True)
    gpu = hfls_amd.Mesh(**kw)
    ora = HTFETI(opts(**kw), kernel_mode="geometric")
    yield request.param, gpu, ora
    gpu.close()


def packed(ora, vecs):
    return np.concatenate(vecs)


def test_setup_kernel_and_fixing_dofs(pair):
    name, gpu, ora = pair
    R, fix, dfc = gpu.kernel()
    assert np.all(dfc == 6)
    off = 0
    for d in range(ora.nsub):
        n = len(ora.subs.l2g[d])
        Rd = R[6 * off:6 * (off + n)].reshape(6, n).T
        off += n
        K = ora.subs.K[d]
        assert np.linalg.norm(K @ Rd) <= 1e-9 * abs(K).max() * np.sqrt(n)
        assert np.allclose(Rd.T @ Rd, np.eye(6), atol=1e-10)
        Q = ora.R[d]
        assert np.linalg.norm(Rd - Q @ (Q.T @ Rd), 2) <= 1e-9
        assert sorted(fix[6 * d:6 * d + 6]) == sorted(ora.fix[d])


def test_load_vector(pair):
    name, gpu, ora = pair
    f = gpu.load_vector()
    fo = packed(ora, ora.subs.f)
    assert np.linalg.norm(f - fo) <= 1e-13 * np.linalg.norm(fo)


def test_operators_match_oracle(pair):
    name, gpu, ora = pair
    rng = np.random.default_rng(11)
    for _ in range(2):
        w = rng.standard_normal(ora.nlam)
        Fw = gpu.mult_Ff(w)
        Fo = ora.mult_Ff(w)
        assert np.linalg.norm(Fw - Fo) <= 1e-9 * np.linalg.norm(Fo)
        Px, _ = gpu.projection(w)
        Po, _ = ora.projection(w)
        assert np.linalg.norm(Px - Po) <= 1e-10 * np.linalg.norm(Po)
        Mx = gpu.preconditioning(w)
        Mo = ora.preconditioning(w)
        assert np.linalg.norm(Mx - Mo) <= 1e-10 * np.linalg.norm(Mo)


def test_kplus_matches_oracle(pair):
    name, gpu, ora = pair
    rng = np.random.default_rng(5)
    fs = [rng.standard_normal(len(l)) for l in ora.subs.l2g]
    x = gpu.mult_Kplus_f(np.concatenate(fs))
    xo = np.concatenate(ora.mult_Kplus_f(fs))
    assert np.linalg.norm(x - xo) <= 1e-9 * np.linalg.norm(xo)


def test_pcpg_matches_oracle_and_direct(pair):
    name, gpu, ora = pair
    eps = 1e-8
    r = gpu.pcpg(eps=eps, max_iter=300)
    ro = ora.pcpg(eps=eps, max_iter=300)
    print(name, "iterations gpu", r["iterations"], "oracle", ro["iterations"], flush=True)
    assert r["converged"]
    assert abs(r["iterations"] - ro["iterations"]) <= 2
    # two iterates stopped at a 1e-8 residual differ by about 1e-8 relative (rounding
    # in the reductions moves them apart at that level): 1e-7 here, and the 1e-8
    # parity bar on the converged solutions below
    assert np.linalg.norm(r["u"] - ro["u"]) <= 1e-7 * np.linalg.norm(ro["u"])
    u_ref, _, _ = global_solve(ora.mesh)
    err = np.linalg.norm(r["u"] - u_ref) / np.linalg.norm(u_ref)
    print(name, "error vs FP64 global direct solve %.3e" % err, flush=True)
    assert err <= 1e-6
    # tight tolerance: displacement within 1e-8 of the oracle's and of the direct solve
    r10 = gpu.pcpg(eps=1e-10, max_iter=500)
    ro10 = ora.pcpg(eps=1e-10, max_iter=500)
    assert r10["converged"]
    e10 = np.linalg.norm(r10["u"] - ro10["u"]) / np.linalg.norm(ro10["u"])
    print(name, "eps 1e-10: error vs oracle %.3e" % e10, flush=True)
    assert e10 <= 1e-8
    assert np.linalg.norm(r10["u"] - u_ref) <= 1e-8 * np.linalg.norm(u_ref)


def test_solve_is_deterministic(pair):
    name, gpu, ora = pair
    a = gpu.pcpg(eps=1e-8)
    b = gpu.pcpg(eps=1e-8)
    assert a["iterations"] == b["iterations"]
    assert np.array_equal(a["u"], b["u"]) and np.array_equal(a["lambda"], b["lambda"])


def test_facade_main_runs():
    """The reference's main() over the C++ facade (hfls_amd/htfeti) converges."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(hfls_amd.LIB_PATH), "htfeti")
    r = subprocess.run([exe, "2", "2", "2", "3", "3", "3"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "converged 1" in r.stdout
