"""GPU tests of the request loop with continuous batching (hfls_server_*).

* every case of every request gives bit-identical displacements, iteration counts
  and residuals whether its request is served alone or together with other
  requests of other meshes arriving while it runs (staggered submissions, fewer
  slots than cases, so cases join free slots mid-flight);
* served results match the FP64 oracle (displacement within 1e-7 relative at
  eps 1e-8, every case converged and certified on its FP64 true residual);
* admission by HBM footprint: with a budget that holds one mesh, a request for a
  second mesh waits until the first is idle, which is then evicted.
"""
import numpy as np
import pytest

import hfls_amd
from oracle.mesh import opts
from oracle.feti import HTFETI

pytestmark = pytest.mark.gpu

MESH_A = dict(Nx=2, Ny=2, Nz=2, nx=4, ny=4, nz=4, bc="clamp_x0", material="uniform", young_modulus="2.1e5",
              load="body")
MESH_B = dict(Cx=2, Cy=1, Cz=1, Nx=2, Ny=2, Nz=2, nx=3, ny=3, nz=4, bc="clamp_x0", material="lognormal",
              young_modulus="2.1e5", seed="3", mat_sigma="1.5", mat_contrast="1e3", load="body")
# (mesh, seed, case0, k)
REQUESTS = [(MESH_A, 11, 0, 3), (MESH_B, 12, 0, 2), (MESH_A, 13, 5, 4), (MESH_B, 14, 2, 1), (MESH_A, 15, 1, 2)]


def serve(reqs, slots, stagger, budget=50.0, precision=32):
    s = hfls_amd.Server(slots=slots, precision=precision, hbm_budget_gb=budget, eps=1e-8, max_iter=500, chunk=4)
    ids, out = [], {}
    pending = list(enumerate(reqs))
    busy = True
    rounds = 0
    while pending or busy:
        for _ in range(1 if stagger else len(pending)):
            if pending:
                i, (m, seed, c0, k) = pending.pop(0)
                ids.append((i, s.submit(m, k=k, seed=seed, case0=c0)))
        busy = s.step()
        rounds += 1
        assert rounds < 2000
    for i, rid in ids:
        assert s.state(rid) == 2
        out[i] = s.result(rid)
    st = s.stats()
    s.close()
    return out, st


@pytest.fixture(scope="module")
def mixed():
    print("serve: mixed run", flush=True)
    return serve(REQUESTS, slots=4, stagger=True)


def test_same_alone_and_in_the_mix(mixed):
    out, st = mixed
    print("mixed stats", st, flush=True)
    assert st["joins"] == sum(r[3] for r in REQUESTS)
    for i, r in enumerate(REQUESTS):
        alone, _ = serve([r], slots=4, stagger=False)
        a, m = alone[0], out[i]
        print("request", i, "iterations alone", a["iterations"], "mixed", m["iterations"], flush=True)
        assert np.array_equal(a["iterations"], m["iterations"])
        assert np.array_equal(a["u"], m["u"])
        assert np.array_equal(a["residual"], m["residual"])


@pytest.mark.parametrize("i", [0, 1])
def test_served_cases_match_oracle(mixed, i):
    from oracle.loads import traction_loads
    out, _ = mixed
    mesh, seed, c0, k = REQUESTS[i]
    ora = HTFETI(opts(**mesh), kernel_mode="geometric")
    F = traction_loads(ora.mesh, ora.subs, seed, range(c0, c0 + k))
    r = out[i]
    assert np.all(r["status"] == 2) and np.all(r["residual"] < 1e-8)
    for j in range(k):
        parts, off = [], 0
        for l2g in ora.subs.l2g:
            parts.append(F[j][off:off + len(l2g)])
            off += len(l2g)
        ro = ora.pcpg(f=parts, eps=1e-8, max_iter=500)
        err = np.linalg.norm(r["u"][j] - ro["u"]) / np.linalg.norm(ro["u"])
        print("request", i, "case", j, "iterations", r["iterations"][j], "oracle", ro["iterations"],
              "err %.2e" % err, flush=True)
        assert err <= 1e-7


def test_admission_by_footprint_and_eviction(mixed):
    out, _ = mixed
    s = hfls_amd.Server(slots=4)
    ea, eb = s.sizes(**MESH_A)[2], s.sizes(**MESH_B)[2]
    s.close()
    budget = 1.05 * max(ea, eb) / 1e9
    assert ea + eb > 1.05 * max(ea, eb)
    reqs = [REQUESTS[0], REQUESTS[1], REQUESTS[2]]
    res, st = serve(reqs, slots=4, stagger=False, budget=budget)
    print("budget %.3f GB, stats" % budget, st, flush=True)
    assert st["evictions"] >= 1 and st["setups"] >= 2
    for i in range(3):
        assert np.array_equal(res[i]["u"], out[i]["u"])
