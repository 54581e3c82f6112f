"""Pins of the oracle's HTFETI chain against a global direct solve and invariants."""
import numpy as np
import pytest

from oracle.mesh import opts
from oracle.feti import HTFETI
from oracle.direct import global_solve

CUBE8_5E3 = dict(Nx=2, Ny=2, Nz=2, nx=5, ny=5, nz=5, bc="clamp_x0", material="uniform",
            young_modulus="2.1e5", load="body", body_force="-1.0")


@pytest.fixture(scope="module")
def cube():
    return HTFETI(opts(eps_iter="1e-8", **CUBE8_5E3))


def test_sizes_cube8_5e3(cube):
    assert cube.nlam == 1878
    assert cube.glue.n_glue == 1446
    assert [len(l) for l in cube.subs.l2g] == [648] * 8
    assert cube.clusters[0]["nlc"] == 72
    assert all(d == 6 for d in cube.defect)


def test_cube8_5e3_matches_direct_solve(cube):
    u_ref, _, _ = global_solve(cube.mesh)
    r = cube.pcpg(eps=1e-11)
    assert np.linalg.norm(r["u"] - u_ref) / np.linalg.norm(u_ref) < 1e-9
    r8 = cube.pcpg(eps=1e-8)
    assert np.linalg.norm(r8["u"] - u_ref) / np.linalg.norm(u_ref) < 1e-7
    assert 8 <= r8["iterations"] <= 20


def test_projector_invariants(cube):
    h = cube
    rng = np.random.default_rng(0)
    x = rng.standard_normal(h.nlam)
    Px, _ = h.projection(x)
    PPx, _ = h.projection(Px)
    assert np.allclose(PPx, Px, atol=1e-12 * np.linalg.norm(x))
    assert np.linalg.norm(h.Gf.T @ Px) <= 1e-11 * np.linalg.norm(x) * np.linalg.norm(h.Gf)
    # G^T lambda0 = e
    e = sum(h.Rf[d].T @ h.subs.f[d] for d in range(h.nsub))
    lam0 = h.Gf @ (h.invGfTGf @ e)
    assert np.allclose(h.Gf.T @ lam0, e, rtol=1e-11, atol=1e-12 * np.abs(e).max())


def test_F_and_M_symmetric(cube):
    h = cube
    rng = np.random.default_rng(1)
    a, b = rng.standard_normal(h.nlam), rng.standard_normal(h.nlam)
    Fa, Fb = h.mult_Ff(a), h.mult_Ff(b)
    assert abs(a @ Fb - b @ Fa) <= 1e-11 * abs(a @ Fa)
    Ma, Mb = h.preconditioning(a), h.preconditioning(b)
    assert abs(a @ Mb - b @ Ma) <= 1e-11 * abs(a @ Ma)
    # F is PSD on range(P)
    Pa, _ = h.projection(a)
    assert Pa @ h.mult_Ff(Pa) > 0


def test_generalised_inverse_invariance(cube):
    """P F P and u do not depend on the s-set / fixing DOFs (reference's random
    rule vs the geometric rule); iteration counts agree within 1."""
    h2 = HTFETI(opts(eps_iter="1e-8", **CUBE8_5E3), kernel_mode="reference", seed=7)
    rng = np.random.default_rng(2)
    x = rng.standard_normal(cube.nlam)
    Px, _ = cube.projection(x)
    y1, _ = cube.projection(cube.mult_Ff(Px))
    y2, _ = h2.projection(h2.mult_Ff(Px))
    assert np.linalg.norm(y1 - y2) <= 1e-10 * np.linalg.norm(y1)
    r1, r2 = cube.pcpg(), h2.pcpg()
    assert abs(r1["iterations"] - r2["iterations"]) <= 1
    assert np.linalg.norm(r1["u"] - r2["u"]) <= 1e-8 * np.linalg.norm(r1["u"])


@pytest.mark.parametrize("kw", [dict(typeBc="cor"), dict(preconditioner="lumped"),
                                dict(preconditioner="Dirichlet_implicit"), dict(Cx=2, Nx=1, Ny=2, Nz=2, nx=3, ny=3, nz=3)])
def test_variants_match_direct(kw):
    k = dict(CUBE8_5E3)
    k.update(nx=3, ny=3, nz=3)
    k.update(kw)
    h = HTFETI(opts(eps_iter="1e-10", **k))
    u_ref, _, _ = global_solve(h.mesh)
    r = h.pcpg()
    assert r["iterations"] < 200
    assert np.linalg.norm(r["u"] - u_ref) / np.linalg.norm(u_ref) < 1e-8
