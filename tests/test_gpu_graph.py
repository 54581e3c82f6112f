"""Captured HIP graphs of the PCPG iteration chunks (plain and block paths).

Every replay must reproduce the eager launches bit for bit: each solve below
replays the captured chunk several times, with the eager initialisation,
recovery and host synchronisations between the replays, and is compared with the
same solve run without graphs.  The block results are also independent of the
host poll interval (a stopped column is frozen on the device).
"""
import numpy as np
import pytest

import hfls_amd

pytestmark = pytest.mark.gpu

CASES = {
    "cube8_5e3": dict(Nx=2, Ny=2, Nz=2, nx=5, ny=5, nz=5, bc="clamp_x0", material="uniform", young_modulus="2.1e5",
                 load="body"),
    "uneven": dict(Cx=2, Cy=1, Cz=1, Nx=2, Ny=2, Nz=2, slabs_x="3,5,4,6", slabs_y="4,3", slabs_z="5,3",
                   bc="clamp_x0", material="lognormal", young_modulus="2.1e5", seed="9", mat_sigma="1.5",
                   mat_contrast="1e3", load="body"),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_graph_replays_match_eager(name):
    kw = CASES[name]
    print("graph case", name, flush=True)
    m = hfls_amd.Mesh(**kw)
    F = m.traction(77, 0, 5)
    m.configure(hip_graph=False)
    p0 = m.pcpg(eps=1e-8, max_iter=500)
    b0 = m.pcpg_block(F, eps=1e-8, max_iter=500, precision=32, want_lambda=True)
    assert p0["converged"] and np.all(b0["status"] == 2)
    m.configure(hip_graph=True)
    for rep in range(3):
        p = m.pcpg(eps=1e-8, max_iter=500)
        assert p["iterations"] == p0["iterations"]
        assert np.array_equal(p["u"], p0["u"]) and np.array_equal(p["lambda"], p0["lambda"])
        b = m.pcpg_block(F, eps=1e-8, max_iter=500, precision=32, want_lambda=True)
        assert np.array_equal(b["iterations"], b0["iterations"])
        assert np.array_equal(b["u"], b0["u"]) and np.array_equal(b["lambda"], b0["lambda"])
    m.configure(poll_chunk=3)
    b = m.pcpg_block(F, eps=1e-8, max_iter=500, precision=32)
    assert np.array_equal(b["u"], b0["u"]) and np.array_equal(b["iterations"], b0["iterations"])
    m.close()


@pytest.mark.parametrize("name", sorted(CASES))
def test_autotuned_gemm_shapes_agree(name):
    """The autotuner times both tile shapes of the row form of the block GEMM on the
    mesh and picks one; either shape gives bit-identical solves."""
    m = hfls_amd.Mesh(**CASES[name])
    m.configure(gemm_kernel="rows")
    F = m.traction(78, 0, 3)
    r_auto = m.pcpg_block(F, eps=1e-8, max_iter=500, precision=32)
    t = m.tuning()
    print(name, "autotune", t, flush=True)
    assert t["row_tiles"] in (1, 2) and t["ms"][1] > 0 and t["ms"][2] > 0 and t["ms"][3] > 0
    out = {}
    for rows in (1, 2):
        m.configure(gemm_rows=rows)
        out[rows] = m.pcpg_block(F, eps=1e-8, max_iter=500, precision=32)
    assert np.array_equal(out[1]["u"], out[2]["u"]) and np.array_equal(out[1]["u"], r_auto["u"])
    m.close()


BAND_CASES = dict(CASES, cube8_8e3=dict(Nx=2, Ny=2, Nz=2, nx=8, ny=8, nz=8, bc="clamp_x0", material="lognormal",
                                        young_modulus="2.1e5", seed="5", load="body"))


@pytest.mark.parametrize("name", sorted(BAND_CASES))
def test_band_gemm_matches_rows(name):
    """The band form of the split-FP16 GEMM (every tile read once, partial sums of the
    bands below a row block added in a fixed order) against the row form: the F and M
    applies on a block of 64 columns within 2e-5 relative of FP64 in both forms and
    within 1e-6 of each other; mixed solves certified with the same iteration counts
    (+-1) and displacements within 1e-7; the band form bitwise reproducible.  cube8_8e3
    has 19 tile rows per subdomain (5 bands, partials from 4 bands below)."""
    kw = BAND_CASES[name]
    print("band case", name, flush=True)
    m = hfls_amd.Mesh(**kw)
    rng = np.random.default_rng(3)
    W = rng.standard_normal((64, m.nlam))
    ref = {op: m.apply_block(op, W, precision=64) for op in (0, 1)}
    res = {}
    for kern in ("rows", "bands"):
        m.configure(gemm_kernel=kern)
        for op in (0, 1):
            y = m.apply_block(op, W, precision=32)
            err = np.linalg.norm(y - ref[op], axis=1) / np.linalg.norm(ref[op], axis=1)
            print(name, kern, "op", op, "max rel err vs FP64 %.2e" % err.max(), flush=True)
            assert err.max() < 2e-5
            res[(kern, op)] = y
    for op in (0, 1):
        d = np.linalg.norm(res[("bands", op)] - res[("rows", op)], axis=1) / np.linalg.norm(ref[op], axis=1)
        assert d.max() < 1e-6
    F = m.traction(81, 0, 4)
    m.configure(gemm_kernel="rows")
    br = m.pcpg_block(F, eps=1e-8, max_iter=800, precision=32)
    m.configure(gemm_kernel="bands")
    bb = m.pcpg_block(F, eps=1e-8, max_iter=800, precision=32)
    bb2 = m.pcpg_block(F, eps=1e-8, max_iter=800, precision=32)
    assert np.all(br["status"] == 2) and np.all(bb["status"] == 2)
    assert np.all(np.abs(bb["iterations"] - br["iterations"]) <= 1)
    du = np.linalg.norm(bb["u"] - br["u"], axis=1) / np.linalg.norm(br["u"], axis=1)
    print(name, "iterations", bb["iterations"], br["iterations"], "u diff %.2e" % du.max(), flush=True)
    assert du.max() < 1e-7
    assert np.array_equal(bb["u"], bb2["u"]) and np.array_equal(bb["iterations"], bb2["iterations"])
    m.close()


def test_measurement_hooks():
    """The timing hook: every kind returns a positive time and work figure, the
    matrix-core ceilings report their flops, the block GEMM its algorithmic flops
    (2 * 64 * sum n_s^2) and HBM bytes (packed 64x64 tiles at 4 B per element, 64
    FP32 columns of right-hand sides and results per interface DOF), the in-solve
    GEMM probe (kind 9, an eager block solve with events around every K+_bb GEMM)
    the same bytes and leaves the captured path bitwise unchanged, and an unknown
    kind is an error."""
    m = hfls_amd.Mesh(**CASES["cube8_5e3"])
    F = m.traction(79, 0, 2)
    r0 = m.pcpg_block(F, eps=1e-8, max_iter=500, precision=32)
    for which in range(10):
        ms, work = m.time_stream(which, reps=2)
        print("hook", which, ms, work, flush=True)
        assert ms > 0 and work > 0
    ms5, fl5 = m.time_stream(5, reps=1)
    assert fl5 == 4.0 * 8 * 256 * 400 * 8 * (2 * 16 * 16 * 4)
    p = hfls_amd.Problem(**CASES["cube8_5e3"])
    ns = [p.sub_info(d)["n_s"] for d in range(p.nsub)]
    tiles = sum(((n + 63) // 64) * ((n + 63) // 64 + 1) // 2 for n in ns)
    assert m.time_stream(2, reps=1)[1] == sum(2.0 * 64 * n * n for n in ns)
    assert m.time_stream(7, reps=1)[1] == 4.0 * 4096 * tiles + 2 * 4.0 * 64 * sum(ns)
    assert m.time_stream(9, reps=4)[1] == m.time_stream(7, reps=1)[1]
    m.configure(gemm_kernel="rows")  # the probe times whichever form is configured
    assert m.time_stream(9, reps=4)[0] > 0
    m.configure(gemm_kernel="bands")
    r1 = m.pcpg_block(F, eps=1e-8, max_iter=500, precision=32)
    assert np.array_equal(r0["u"], r1["u"]) and np.array_equal(r0["iterations"], r1["iterations"])
    with pytest.raises(hfls_amd.HflsError):
        m.time_stream(10, reps=1)
    m.close()
