"""Failure paths through the C ABI on the GPU.  Every case is forced from the host
(bad loads, tiny iteration limits, option combinations that fail in the host part
of the set-up, cancellation); every kernel still runs to its normal end.

* a NaN load column of a block ends alone with status 6 (non-finite residual);
  the other columns are bit-identical to the same block without it;
* the plain solve with a NaN load raises HflsError;
* the iteration limit: converged False (plain), status 3 (block);
* the server: a request whose mesh set-up fails ends with status 7 and the error
  message, later requests are served; a cancelled request leaves its slots and
  the others finish as they would alone.
"""
import numpy as np
import pytest

import hfls_amd

pytestmark = pytest.mark.gpu

SMALL = dict(Nx=2, Ny=2, Nz=2, nx=3, ny=3, nz=3, bc="clamp_x0", material="uniform", young_modulus="2.1e5",
             load="body")
SEED = 20240611


@pytest.fixture(scope="module")
def mesh():
    print("failure paths on", SMALL, flush=True)
    m = hfls_amd.Mesh(**SMALL)
    yield m
    m.close()


@pytest.mark.parametrize("precision", [64, 32])
def test_nan_column_stops_alone(mesh, precision):
    F = mesh.traction(SEED, 0, 4)
    clean = mesh.pcpg_block(F, eps=1e-8, max_iter=300, precision=precision)
    Fb = F.copy()
    Fb[2, 5] = np.nan
    r = mesh.pcpg_block(Fb, eps=1e-8, max_iter=300, precision=precision)
    print("precision", precision, "status", r["status"], "iterations", r["iterations"], flush=True)
    assert r["status"][2] == 6
    for j in (0, 1, 3):
        assert r["status"][j] == 2 and r["iterations"][j] == clean["iterations"][j]
        assert np.array_equal(r["u"][j], clean["u"][j])


def test_nan_column_concurrent_blocks(mesh):
    F = mesh.traction(SEED, 0, 4)
    F[1, 0] = np.inf
    r = mesh.pcpg_blocks(F, k=2, eps=1e-8, max_iter=300, precision=32)
    print("blocks status", r["status"], flush=True)
    assert r["status"][1] == 6 and np.all(r["status"][[0, 2, 3]] == 2)


def test_plain_nan_raises(mesh):
    f = mesh.load_vector()
    f[3] = np.nan
    with pytest.raises(hfls_amd.HflsError, match="non-finite"):
        mesh.pcpg(eps=1e-8, max_iter=100, f_packed=f)
    # the mesh is still usable
    r = mesh.pcpg(eps=1e-8, max_iter=100)
    assert r["converged"]


def test_iteration_limit(mesh):
    r = mesh.pcpg(eps=1e-14, max_iter=2)
    assert not r["converged"] and r["iterations"] == 2
    b = mesh.pcpg_block(mesh.traction(SEED, 0, 2), eps=1e-14, max_iter=3, precision=64)
    assert np.all(b["status"] == 3) and np.all(b["iterations"] == 3)


def test_server_failed_setup_then_serves():
    # corner plus averaged-edge gluing with every neighbour pair (Bc_fullRank=false)
    # on 3x3x3 subdomains puts 276 cluster constraints on the centre subdomain, more
    # than the set-up handles (256): the host part of the set-up throws after the
    # factorisation has finished on the device
    bad = dict(SMALL, Nx=3, Ny=3, Nz=3, nx=2, ny=2, nz=2, typeBc="cor+edge", Bc_fullRank="false")
    srv = hfls_amd.Server(slots=8, precision=32, eps=1e-8, max_iter=300, chunk=4, hbm_budget_gb=50.0)
    try:
        a = srv.submit(bad, k=2, seed=SEED)
        b = srv.submit(SMALL, k=3, seed=SEED)
        while srv.step():
            pass
        with pytest.raises(hfls_amd.HflsError, match="set-up of the mesh failed") as ei:
            srv.result(a)
        assert np.all(ei.value.status == 7)
        rb = srv.result(b)
        assert np.all(rb["status"] == 2)
        st = srv.stats()
        print("server stats", st, flush=True)
        assert st["failed_setups"] == 1
        # a later request on the good mesh is served from the cache
        c = srv.submit(SMALL, k=1, seed=SEED, case0=5)
        while srv.step():
            pass
        assert srv.result(c)["status"][0] == 2
    finally:
        srv.close()


def test_server_cancel():
    srv = hfls_amd.Server(slots=4, precision=32, eps=1e-8, max_iter=300, chunk=2, hbm_budget_gb=50.0)
    alone = hfls_amd.Server(slots=4, precision=32, eps=1e-8, max_iter=300, chunk=2, hbm_budget_gb=50.0)
    try:
        a = srv.submit(SMALL, k=3, seed=SEED, case0=0)
        b = srv.submit(SMALL, k=2, seed=SEED, case0=10)
        srv.step()
        srv.step()
        srv.cancel(a)
        assert srv.state(a) == -1
        while srv.step():
            pass
        rb = srv.result(b)
        c = alone.submit(SMALL, k=2, seed=SEED, case0=10)
        while alone.step():
            pass
        rc = alone.result(c)
        assert np.all(rb["status"] == 2)
        assert np.array_equal(rb["u"], rc["u"])
        assert srv.stats()["cancelled"] == 1
    finally:
        srv.close()
        alone.close()
