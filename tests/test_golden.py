"""CPU checks against the committed golden fixtures (tests/golden/golden.py wrote
them from the oracle).

* the 2x2x2 x 5^3 cube: the oracle and the product's host library both reproduce the full
  integer arrays of the bit-exact contract (Dirichlet/corner DOFs, Bf row table and
  weights, l2g, compressed Bf, j_col_cmpr) bit for bit;
* the 512-subdomain cube: the product's arrays have the fixture's SHA-256
  digests;
* the 2x2x2 x 10^3 displacement fixture agrees with the global FP64 direct solve
  (1e-8 relative), which pins the fixture itself.
"""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)
import golden  # noqa: E402


def _fixture0():
    with np.load(os.path.join(HERE, "cube8_5e3_integer.npz")) as f:
        return {k: f[k] for k in golden.KEYS}


@pytest.mark.parametrize("source", ["oracle", "native"])
def test_cube8_5e3_integer_arrays(source):
    ref = _fixture0()
    got = golden.oracle_arrays(golden.CUBE8_5E3) if source == "oracle" else golden.native_arrays(golden.CUBE8_5E3)
    assert len(ref["weight"]) == 1878  # 1,446 glue rows + 432 Dirichlet rows
    for k in golden.KEYS:
        assert got[k].dtype.kind == ref[k].dtype.kind, k
        assert np.array_equal(got[k], ref[k]), k


def test_cube512_10e3_native_digests():
    with open(os.path.join(HERE, "cube512_10e3_integer_sha256.json")) as f:
        meta = json.load(f)
    got = golden.native_arrays(meta["options"])
    assert len(got["weight"]) == meta["nlam"] == 588846 and len(got["l2g_ptr"]) - 1 == meta["nsub"] == 512
    assert golden.digests(got) == meta["sha256"]


def test_cube8_displacement_fixture_matches_direct_solve():
    from oracle.mesh import Mesh, opts
    from oracle.direct import global_solve
    with np.load(os.path.join(HERE, "cube8_10e3_u.npz")) as f:
        u = f["u"]
        it8 = int(f["iterations_1e8"])
    m = Mesh(opts(**golden.CUBE8_10E3))
    u_ref, _, _ = global_solve(m)
    assert u.shape == u_ref.shape == (3 * 21 ** 3,)
    assert np.linalg.norm(u - u_ref) <= 1e-8 * np.linalg.norm(u_ref)
    assert 5 <= it8 <= 100
