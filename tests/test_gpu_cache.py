"""GPU tests of the mesh cache: saved set-up state in a file (hfls_mesh_save /
hfls_mesh_restore) and the serving loop's pinned-host-memory cache of evicted
meshes (hfls_server_host_cache).

* a restored mesh is bit-identical to the mesh that was saved: kernels, fixing
  DOFs, loads, every operator of the C ABI on seeded inputs, the plain FP64 solve
  and a mixed-precision block of face-traction cases; saving the restored mesh
  writes the same bytes again;
* a file whose options, sizes or indices do not describe the mesh is rejected on
  the host before any kernel reads it (the fingerprint, element counts, constraint
  rows and sources, fixing DOFs, truncation), and the intact file still restores;
* a served mesh evicted for lack of HBM is restored from host memory when it is
  requested again, with the same displacements as its first run and as a server
  without the host cache; a host cache too small for a mesh falls back to set-ups.
"""
import struct

import numpy as np
import pytest

import hfls_amd

pytestmark = pytest.mark.gpu

MULTI = dict(Cx=2, Cy=1, Cz=1, Nx=2, Ny=2, Nz=2, nx=3, ny=3, nz=4, bc="clamp_x0", material="lognormal",
             young_modulus="2.1e5", seed="3", mat_sigma="0.5", load="body")
CASES = {
    "multi_cor": dict(MULTI, typeBc="cor"),
    "cube_lumped": dict(Nx=2, Ny=2, Nz=2, nx=5, ny=5, nz=5, bc="clamp_x0", material="uniform", young_modulus="2.1e5",
                        load="body", preconditioner="lumped"),
    "uneven_ac_regularized": dict(Cx=2, Cy=1, Cz=1, Nx=2, Ny=2, Nz=2, slab_choices="2,3,4", bc="clamp_x0",
                                  material="lognormal", young_modulus="2.1e5", seed="5", mat_sigma="1.5",
                                  mat_contrast="1e3", load="body", Ac_extended_by_kerGc="false"),
    "multi_cor_edge": dict(MULTI, typeBc="cor+edge"),
}


def fields(b):
    """Offsets of the saved fields: name -> (data offset, element size, count)."""
    off = 12
    (nopt,) = struct.unpack_from("<I", b, off)
    off += 4
    for _ in range(2 * nopt):
        (n,) = struct.unpack_from("<I", b, off)
        off += 4 + n
    out = {}
    while b[off:off + 8] != b"HFLS_END":
        name = b[off:off + 16].rstrip(b"\0").decode()
        elem, count = struct.unpack_from("<IQ", b, off + 16)
        out[name] = (off + 28, elem, count)
        off += 28 + elem * count
    assert off + 8 == len(b)
    return out


@pytest.fixture(scope="module", params=sorted(CASES))
def saved(request, tmp_path_factory):
    name = request.param
    print("cache case", name, flush=True)
    path = tmp_path_factory.mktemp("cache") / (name + ".mesh")
    m = hfls_amd.Mesh(**CASES[name])
    m.save(str(path))
    r = hfls_amd.Mesh.restore(str(path))
    print("set-up %.1f ms, restore %.1f ms, file %d bytes" % (m.setup_ms, r.restore_ms, path.stat().st_size),
          flush=True)
    yield name, m, r, path
    r.close()
    m.close()


def test_restored_mesh_is_bit_identical(saved):
    name, m, r, path = saved
    assert (r.nsub, r.nlam, r.ncol, r.npacked, r.nPoints) == (m.nsub, m.nlam, m.ncol, m.npacked, m.nPoints)
    assert r.setup_ms == m.setup_ms and r.restore_ms > 0
    for a, b in zip(m.kernel(), r.kernel()):
        assert np.array_equal(a, b)
    assert np.array_equal(m.load_vector(), r.load_vector())
    assert np.array_equal(m.Rf(), r.Rf())
    assert m.counts() == r.counts()
    rng = np.random.default_rng(7)
    w = rng.standard_normal(m.nlam)
    f = rng.standard_normal(m.npacked)
    a = rng.standard_normal(m.ncol)
    assert np.array_equal(m.mult_Ff(w), r.mult_Ff(w))
    for x, y in zip(m.projection(w), r.projection(w)):
        assert np.array_equal(x, y)
    assert np.array_equal(m.preconditioning(w), r.preconditioning(w))
    assert np.array_equal(m.mult_Kplus_f(f), r.mult_Kplus_f(f))
    assert np.array_equal(m.mult_Gf(a), r.mult_Gf(a))
    assert np.array_equal(m.mult_GfT(w), r.mult_GfT(w))
    assert np.array_equal(m.mult_RfT(f), r.mult_RfT(f))
    p1, p2 = m.pcpg(eps=1e-8, max_iter=500), r.pcpg(eps=1e-8, max_iter=500)
    print(name, "plain iterations", p1["iterations"], p2["iterations"], flush=True)
    assert p1["converged"] and p1["iterations"] == p2["iterations"]
    assert np.array_equal(p1["u"], p2["u"]) and np.array_equal(p1["lambda"], p2["lambda"])
    F = m.traction(21, 0, 8)
    b1 = m.pcpg_block(F, eps=1e-8, max_iter=500, precision=32)
    b2 = r.pcpg_block(F, eps=1e-8, max_iter=500, precision=32)
    print(name, "block iterations", b1["iterations"], b2["iterations"], flush=True)
    assert np.all(b1["status"] == 2)
    assert np.array_equal(b1["iterations"], b2["iterations"]) and np.array_equal(b1["u"], b2["u"])
    again = path.with_suffix(".again")
    r.save(str(again))
    assert again.read_bytes() == path.read_bytes()


def _rejects(path, data, message):
    path.write_bytes(data)
    with pytest.raises(hfls_amd.HflsError) as e:
        hfls_amd.Mesh.restore(str(path))
    print("rejected:", e.value, flush=True)
    assert message in str(e.value)


def test_restore_checks_the_saved_state(saved, tmp_path):
    name, m, _, path = saved
    b = path.read_bytes()
    fl = fields(b)
    bad = tmp_path / "bad.mesh"
    # another mesh: the options say nx = 4 where the state was saved with 3 (or 5)
    nx = CASES[name].get("nx")
    if nx is not None:
        key = struct.pack("<I", 2) + b"nx" + struct.pack("<I", 1) + str(nx).encode()
        assert b.count(key) == 1
        _rejects(bad, b.replace(key, key[:-1] + str(nx + 1).encode()), "does not match")
    # fingerprint: nsub
    off = fl["fingerprint"][0]
    c = bytearray(b)
    struct.pack_into("<q", c, off, m.nsub + 1)
    _rejects(bad, bytes(c), "does not match")
    # a constraint row outside its cluster, a constraint source past the rows, a fixing
    # DOF outside the interface
    for fname, value, message in (("lcmap", 1 << 30, "constraint row outside its cluster"),
                                  ("gcsrc", 1 << 30, "constraint source out of range"),
                                  ("fixpos", -1, "fixing DOF outside")):
        off, elem, count = fl[fname]
        assert elem == 4 and count > 0
        c = bytearray(b)
        struct.pack_into("<i", c, off + 4 * (count // 2), value)
        _rejects(bad, bytes(c), message)
    # a count that differs from the mesh's layout
    off, elem, count = fl["Kt"]
    c = bytearray(b)
    struct.pack_into("<Q", c, off - 8, count - 1)
    _rejects(bad, bytes(c), "entries, the mesh of its options needs")
    # truncated inside the tiles and before the end marker
    _rejects(bad, b[:off + elem * count // 2], "truncated")
    _rejects(bad, b[:-4], "truncated")
    # the intact file still restores, and solves as the saved mesh does
    r = hfls_amd.Mesh.restore(str(path))
    assert np.array_equal(r.pcpg(eps=1e-8, max_iter=500)["u"], m.pcpg(eps=1e-8, max_iter=500)["u"])
    r.close()


MESH_A = dict(Nx=2, Ny=2, Nz=2, nx=4, ny=4, nz=4, bc="clamp_x0", material="uniform", young_modulus="2.1e5",
              load="body")
MESH_B = dict(Cx=2, Cy=1, Cz=1, Nx=2, Ny=2, Nz=2, nx=3, ny=3, nz=4, bc="clamp_x0", material="lognormal",
              young_modulus="2.1e5", seed="3", mat_sigma="1.5", mat_contrast="1e3", load="body")
# A, then B (A is evicted), then A again (B is evicted, A restored or set up again)
STREAM = [(MESH_A, 11, 0, 3), (MESH_B, 12, 0, 2), (MESH_A, 11, 0, 3)]


def serve_sequence(host_cache_gb):
    s = hfls_amd.Server(slots=4, precision=32, eps=1e-8, max_iter=500, chunk=4)
    ea, eb = s.sizes(**MESH_A)[2], s.sizes(**MESH_B)[2]
    s.close()
    budget = 1.05 * max(ea, eb) / 1e9
    s = hfls_amd.Server(slots=4, precision=32, eps=1e-8, max_iter=500, chunk=4, hbm_budget_gb=budget,
                        host_cache_gb=host_cache_gb)
    out = []
    for mesh, seed, c0, k in STREAM:
        rid = s.submit(mesh, k=k, seed=seed, case0=c0)
        rounds = 0
        while s.step():
            rounds += 1
            assert rounds < 2000
        out.append(s.result(rid))
    st = s.stats()
    s.close()
    print("host cache %.3g GB:" % host_cache_gb, {k: st[k] for k in ("setups", "evictions", "saves", "restores",
                                                                       "host_drops", "setup_ms", "restore_ms",
                                                                       "save_ms", "host_bytes")}, flush=True)
    return out, st


def test_server_restores_evicted_mesh_from_host_memory():
    plain, st0 = serve_sequence(0.0)
    assert st0["setups"] == 3 and st0["restores"] == 0 and st0["saves"] == 0
    cached, st1 = serve_sequence(20.0)
    assert st1["setups"] == 2 and st1["restores"] == 1 and st1["saves"] == 2 and st1["evictions"] == 2
    assert st1["host_meshes"] == 2 and st1["host_bytes"] > 0
    tiny, st2 = serve_sequence(1e-6)  # nothing fits: every eviction falls back to a later set-up
    assert st2["setups"] == 3 and st2["restores"] == 0 and st2["host_meshes"] == 0
    for runs in (plain, cached, tiny):
        assert all(np.all(r["status"] == 2) for r in runs)
        assert np.array_equal(runs[2]["u"], plain[0]["u"])
        assert np.array_equal(runs[2]["iterations"], plain[0]["iterations"])
        assert np.array_equal(runs[1]["u"], plain[1]["u"])
