"""Edge cases of the C ABI and its Python mirror on the GPU path.

  * a zero load is already the solution: u = 0 with no iteration, in the plain
    path and as one column of a block, whose other columns are unaffected
    (bit-identical to solving them alone)
  * the largest block (k = 64) gives each case the same bits as solving it alone
  * k outside 1..64 and arrays of the wrong shape are errors, never reads past
    the caller's buffers
"""
import numpy as np
import pytest

import hfls_amd

pytestmark = pytest.mark.gpu

CUBE8_5E3 = dict(Nx=2, Ny=2, Nz=2, nx=5, ny=5, nz=5, bc="clamp_x0", material="uniform",
            young_modulus="2.1e5", load="body", body_force="-1.0")


@pytest.fixture(scope="module")
def mesh():
    print("edge case mesh", CUBE8_5E3, flush=True)
    m = hfls_amd.Mesh(**CUBE8_5E3)
    yield m
    m.close()


def test_zero_load_plain(mesh):
    r = mesh.pcpg(eps=1e-8, max_iter=500, f_packed=np.zeros(mesh.npacked))
    assert r["iterations"] == 0 and r["converged"]
    assert np.all(r["u"] == 0.0) and np.all(r["lambda"] == 0.0)


@pytest.mark.parametrize("precision", [64, 32])
def test_zero_column_in_block(mesh, precision):
    F = mesh.traction(31, 0, 3)
    F[1] = 0.0
    b = mesh.pcpg_block(F, eps=1e-8, max_iter=500, precision=precision)
    assert b["iterations"][1] == 0 and b["status"][1] == 2 and np.all(b["u"][1] == 0.0)
    assert np.all(b["status"] == 2)
    for j in (0, 2):
        a = mesh.pcpg_block(F[j:j + 1], eps=1e-8, max_iter=500, precision=precision)
        assert np.array_equal(a["u"][0], b["u"][j]) and a["iterations"][0] == b["iterations"][j]


def test_largest_block_same_as_alone(mesh):
    F = mesh.traction(32, 0, 64)
    b = mesh.pcpg_block(F, eps=1e-8, max_iter=500, precision=32)
    assert np.all(b["status"] == 2) and np.all(b["residual"] < 1e-8)
    for j in (0, 37, 63):
        a = mesh.pcpg_block(F[j:j + 1], eps=1e-8, max_iter=500, precision=32)
        assert np.array_equal(a["u"][0], b["u"][j]) and a["iterations"][0] == b["iterations"][j]


def test_invalid_sizes_are_errors(mesh):
    F = mesh.traction(33, 0, 65)
    with pytest.raises(hfls_amd.HflsError):
        mesh.pcpg_block(F, eps=1e-8, max_iter=50)
    with pytest.raises(hfls_amd.HflsError):
        mesh.pcpg_block(k=0, eps=1e-8, max_iter=50)
    with pytest.raises(ValueError):
        mesh.pcpg_block(np.zeros((0, mesh.npacked)))
    with pytest.raises(ValueError):
        mesh.pcpg_block(np.zeros((2, mesh.npacked - 1)))
    with pytest.raises(ValueError):
        mesh.pcpg(f_packed=np.zeros(mesh.npacked + 1))
    with pytest.raises(ValueError):
        mesh.mult_Ff(np.zeros(mesh.nlam - 1))
    with pytest.raises(ValueError):
        mesh.pcpg_blocks(np.zeros((4, mesh.npacked)), k=3)
