"""GPU parity of the block (multi-load-case) PCPG through the C ABI.

Each column of a block is one load case run through the reference iteration
(Cluster.cpp:1363-1458) with its own scalars.  Tolerances:
  * F applied to a block, FP64: 1e-9 relative to the oracle per column;
    FP32 operator (precision 32): 2e-5 relative (FP32 storage and accumulation of
    K+_bb; the cluster correction stays FP64)
  * block FP64 solve: iterations within +-2 of the oracle at eps 1e-8, displacement
    within 1e-8 relative of the oracle's with both solved to eps 1e-10
  * block mixed-precision solve at eps 1e-8: every case converged with its FP64
    true residual below eps, displacement within 1e-7 relative of the FP64 oracle
    and error vs the FP64 global direct solve printed per case
  * a case gives bit-identical results alone (k = 1) and inside a block
"""
import numpy as np
import pytest

import hfls_amd
from oracle.mesh import opts
from oracle.feti import HTFETI
from oracle.direct import global_solve
from oracle.loads import traction_loads

pytestmark = pytest.mark.gpu

CUBE8_5E3 = dict(Nx=2, Ny=2, Nz=2, nx=5, ny=5, nz=5, bc="clamp_x0", material="uniform",
            young_modulus="2.1e5", load="body", body_force="-1.0")
MULTI = dict(Cx=2, Cy=1, Cz=1, Nx=2, Ny=2, Nz=2, nx=3, ny=3, nz=4, bc="clamp_x0", material="lognormal",
             young_modulus="2.1e5", seed="3", mat_sigma="0.5", load="body")
UNEVEN = dict(Cx=2, Cy=1, Cz=1, Nx=2, Ny=2, Nz=2, slab_choices="2,3,4", bc="clamp_x0", material="lognormal",
              young_modulus="2.1e5", seed="5", mat_sigma="1.5", mat_contrast="1e3", load="body")
CASES = {"cube8_5e3": CUBE8_5E3, "multi": MULTI, "uneven": UNEVEN,
         # the other gluing and preconditioner of the reference (Cluster.cpp:846-903, 1106-1107)
         "cor": dict(MULTI, typeBc="cor"), "lumped": dict(CUBE8_5E3, preconditioner="lumped"),
         "cube8_5e3_cor_edge": dict(CUBE8_5E3, typeBc="cor+edge")}
SEED = 20240611
K = 5


@pytest.fixture(scope="module", params=sorted(CASES))
def pair(request):
    kw = CASES[request.param]
    print("case", request.param, kw, flush=True)
    gpu = hfls_amd.Mesh(**kw)
    ora = HTFETI(opts(**kw), kernel_mode="geometric")
    F = gpu.traction(SEED, 0, K)
    F[K - 1] = gpu.load_vector()  # last column: the body force of the options
    yield request.param, gpu, ora, F
    gpu.close()


def split(ora, fpacked):
    out, off = [], 0
    for l2g in ora.subs.l2g:
        out.append(fpacked[off:off + len(l2g)])
        off += len(l2g)
    return out


def test_traction_matches_oracle(pair):
    name, gpu, ora, F = pair
    Fo = traction_loads(ora.mesh, ora.subs, SEED, range(K - 1))
    for j in range(K - 1):
        assert np.array_equal(F[j], Fo[j])


def test_block_F_fp64_and_fp32(pair):
    name, gpu, ora, F = pair
    rng = np.random.default_rng(7)
    W = rng.standard_normal((4, ora.nlam))
    F64 = gpu.mult_Ff_block(W, precision=64)
    F32 = gpu.mult_Ff_block(W, precision=32)
    for j in range(4):
        Fo = ora.mult_Ff(W[j])
        e64 = np.linalg.norm(F64[j] - Fo) / np.linalg.norm(Fo)
        e32 = np.linalg.norm(F32[j] - Fo) / np.linalg.norm(Fo)
        print(name, "F block column", j, "fp64 %.2e fp32 %.2e" % (e64, e32), flush=True)
        assert e64 <= 1e-9
        assert e32 <= 2e-5


@pytest.mark.parametrize("fmt", ["f16x2", "f32"])
def test_matrix_core_formats(fmt):
    """K+_bb and S_i of the mixed-precision operators in both matrix-core formats
    (block_gemm: split FP16 hi + lo planes with power-of-two scales, or FP32): F and M at
    precision 32 within 2e-5 of FP64 per column, also for columns scaled by 1e-12 and 1e6
    (per-column scales of the right-hand sides) and, for M, 1e-35, and a mixed solve to eps 1e-8 within
    1e-7 of the FP64 solve at eps 1e-10."""
    kw = dict(UNEVEN, block_gemm=fmt)
    print("block_gemm", fmt, kw, flush=True)
    gpu = hfls_amd.Mesh(**kw)
    try:
        rng = np.random.default_rng(11)
        W = rng.standard_normal((4, gpu.nlam))
        W[1] *= 1e-12
        W[2] *= 1e6
        for op in (0, 1):
            A64 = gpu.apply_block(op, W, 64)
            A32 = gpu.apply_block(op, W, 32)
            for j in range(4):
                e = np.linalg.norm(A32[j] - A64[j]) / np.linalg.norm(A64[j])
                print(fmt, "op", op, "column", j, "rel %.2e" % e, flush=True)
                assert e <= 2e-5
        # a column below 2^-111 (smallest power-of-two scale); through M its results
        # stay in the FP32 normal range
        Wt = W[:1] * 1e-35
        e = np.linalg.norm(gpu.apply_block(1, Wt, 32)[0] - gpu.apply_block(1, Wt, 64)[0])
        e /= np.linalg.norm(gpu.apply_block(1, Wt, 64)[0])
        print(fmt, "op 1 column scaled 1e-35 rel %.2e" % e, flush=True)
        assert e <= 2e-5
        F = gpu.traction(SEED, 0, 3)
        r = gpu.pcpg_block(F, eps=1e-8, max_iter=300, precision=32)
        r64 = gpu.pcpg_block(F, eps=1e-10, max_iter=500, precision=64)
        for j in range(3):
            e = np.linalg.norm(r["u"][j] - r64["u"][j]) / np.linalg.norm(r64["u"][j])
            print(fmt, "mixed case", j, "iterations", r["iterations"][j], "rel %.2e" % e, flush=True)
            assert r["status"][j] == 2 and e <= 1e-7
    finally:
        gpu.close()


def test_block_fp64_matches_oracle(pair):
    """Iterations at eps 1e-8 within +-2 of the oracle's; displacements compared at
    eps 1e-10, where both solves are converged well below the 1e-8 bar (at 1e-8 the
    corner-only gluing's weaker coarse space leaves the two displacements up to
    ~2e-8 apart after 41 and 42 iterations)."""
    name, gpu, ora, F = pair
    r = gpu.pcpg_block(F, eps=1e-8, max_iter=300, precision=64)
    r10 = gpu.pcpg_block(F, eps=1e-10, max_iter=500, precision=64)
    assert np.all(r["status"] == 2) and np.all(r10["status"] == 2)
    for j in range(K):
        ro = ora.pcpg(f=split(ora, F[j]), eps=1e-8, max_iter=300)
        ro10 = ora.pcpg(f=split(ora, F[j]), eps=1e-10, max_iter=500)
        err = np.linalg.norm(r10["u"][j] - ro10["u"]) / np.linalg.norm(ro10["u"])
        print(name, "fp64 case", j, "iterations", r["iterations"][j], "oracle", ro["iterations"],
              "err at 1e-10 %.2e" % err, flush=True)
        assert abs(int(r["iterations"][j]) - ro["iterations"]) <= 2
        assert err <= 1e-8


def test_block_mixed_precision(pair):
    name, gpu, ora, F = pair
    eps = 1e-8
    r = gpu.pcpg_block(F, eps=eps, max_iter=300, precision=32)
    print(name, "mixed: sweeps", r["sweeps"], "certifications", r["cert_steps"], "iterations", r["iterations"],
          "residuals", r["residual"], flush=True)
    assert np.all(r["status"] == 2)
    assert np.all(r["residual"] < eps)
    for j in range(K):
        ro = ora.pcpg(f=split(ora, F[j]), eps=eps, max_iter=300)
        err_o = np.linalg.norm(r["u"][j] - ro["u"]) / np.linalg.norm(ro["u"])
        u_ref, _, _ = global_solve(ora.mesh, split(ora, F[j]), ora.subs)
        err_d = np.linalg.norm(r["u"][j] - u_ref) / np.linalg.norm(u_ref)
        print(name, "mixed case", j, "iterations", r["iterations"][j], "oracle", ro["iterations"],
              "err vs oracle %.2e vs direct %.2e" % (err_o, err_d), flush=True)
        assert err_o <= 1e-7


@pytest.mark.parametrize("precision", [64, 32])
def test_same_alone_and_in_block(pair, precision):
    name, gpu, ora, F = pair
    blk = gpu.pcpg_block(F, eps=1e-8, max_iter=300, precision=precision, want_lambda=True)
    for j in (0, K - 1):
        one = gpu.pcpg_block(F[j:j + 1], eps=1e-8, max_iter=300, precision=precision, want_lambda=True)
        assert one["iterations"][0] == blk["iterations"][j]
        assert np.array_equal(one["u"][0], blk["u"][j])
        assert np.array_equal(one["lambda"][0], blk["lambda"][j])
    # reordered block: case 0 moved to the last column
    perm = list(range(1, K)) + [0]
    rb = gpu.pcpg_block(F[perm], eps=1e-8, max_iter=300, precision=precision)
    assert np.array_equal(rb["u"][K - 1], blk["u"][0])


def test_block_into_pinned_buffers(pair):
    """Loads in and displacements out of page-locked host memory (hfls_host_alloc),
    the out buffer reused over two solves: bit-identical to pageable arrays; a
    wrongly shaped out array is rejected before the solve."""
    name, gpu, ora, F = pair
    ref = gpu.pcpg_block(F, eps=1e-8, max_iter=300, precision=32)
    Fp = hfls_amd.pinned_empty(F.shape)
    Fp[:] = F
    up = hfls_amd.pinned_empty((K, 3 * gpu.nPoints))
    for rep in range(2):
        up[:] = np.nan
        r = gpu.pcpg_block(Fp, eps=1e-8, max_iter=300, precision=32, out=up)
        assert r["u"] is up
        assert np.array_equal(up, ref["u"])
        assert np.array_equal(r["iterations"], ref["iterations"])
    with pytest.raises(ValueError):
        gpu.pcpg_block(Fp, eps=1e-8, max_iter=300, precision=32, out=up[:, :-1])
    del Fp, up


@pytest.mark.parametrize("nblocks", [2, 3])
def test_concurrent_blocks_same_as_single(pair, nblocks):
    """Blocks solved concurrently in separate block contexts (own buffers and
    streams, shared mesh operators) give bit-identical results to one block at a
    time."""
    name, gpu, ora, F = pair
    one = gpu.pcpg_block(F, eps=1e-8, max_iter=300, precision=32)
    order = [F, F[::-1], F[1:].tolist() + [F[0]]][:nblocks]
    Fs = np.concatenate([np.asarray(f) for f in order])
    r = gpu.pcpg_blocks(Fs, k=K, eps=1e-8, max_iter=300, precision=32)
    print(name, "blocks", nblocks, "span %.1f ms" % r["solve_ms"], flush=True)
    assert np.all(r["status"] == 2)
    perms = [list(range(K)), list(range(K))[::-1], list(range(1, K)) + [0]][:nblocks]
    for b, p in enumerate(perms):
        for j, src in enumerate(p):
            assert np.array_equal(r["u"][b * K + j], one["u"][src])
            assert r["iterations"][b * K + j] == one["iterations"][src]
