"""Golden fixtures of the bit-exact integer contract and of a mid-size solve.

Run from the repository root to (re)write the fixtures next to this file:

    python tests/golden/golden.py

Written from the oracle (oracle/, the CPU restatement of the reference, pinned by
its own tests against closed forms and a global direct solve):

* ``cube8_5e3_integer.npz`` -- the 2x2x2 cube of 5^3-element subdomains (x = 0 clamped):
  Dirichlet and corner DOFs, the Bf row table (Cluster.cpp:846-924: subdomain and
  global DOF of each row's +1 and -1 entry, weights), per-subdomain l2g
  (Data.cpp:111-144), compressed Bf (rows, local columns, signs;
  Cluster.cpp:931-959) and j_col_cmpr (Cluster.cpp:756-765), as full arrays.
* ``cube512_10e3_integer_sha256.json`` -- the same arrays for the 512-subdomain cube (8 clusters x 4^3
  subdomains of 10^3 hex8), as SHA-256 digests of their little-endian bytes.
* ``cube8_10e3_u.npz`` -- displacement of the 2x2x2 x 10^3 cube (x = 0 clamped,
  uniform E, body force) from the oracle's PCPG at eps 1e-10 (sparse LU K_reg,
  implicit K+), with the iteration count at eps 1e-8.

The arrays are built by ``oracle_arrays`` (from the oracle) and ``native_arrays``
(from the product's host library, hfls_amd.Problem) in one canonical layout, so a
test compares them array by array or digest by digest.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

CUBE8_5E3 = dict(Nx=2, Ny=2, Nz=2, nx=5, ny=5, nz=5, bc="clamp_x0", material="uniform", young_modulus="2.1e5",
            load="body", body_force="-1.0")
CUBE512_10E3 = dict(Cx=2, Cy=2, Cz=2, Nx=4, Ny=4, Nz=4, nx=10, ny=10, nz=10, bc="clamp_x0")
CUBE8_10E3 = dict(Nx=2, Ny=2, Nz=2, nx=10, ny=10, nz=10, bc="clamp_x0", material="uniform", young_modulus="2.1e5",
             load="body", body_force="-1.0")
KEYS = ["dirichlet", "corners", "row_a_sub", "row_a_dof", "row_b_sub", "weight", "l2g", "l2g_ptr", "bf_rows",
        "bf_col", "bf_sign", "bf_ptr", "jcol", "jcol_ptr"]


def _concat(parts):
    ptr = np.zeros(len(parts) + 1, dtype=np.int64)
    ptr[1:] = np.cumsum([len(p) for p in parts])
    flat = np.concatenate([np.asarray(p, dtype=np.int64) for p in parts]) if parts else np.zeros(0, np.int64)
    return flat, ptr


def oracle_arrays(kw):
    from oracle.mesh import Mesh, opts
    from oracle.fe import Subdomains
    from oracle.decomp import Gluing
    m = Mesh(opts(**kw))
    s = Subdomains(m)
    g = Gluing(m, s)
    out = {
        "dirichlet": np.asarray(m.dirichlet_dofs, dtype=np.int64),
        "corners": np.asarray(m.corner_dofs, dtype=np.int64),
        "row_a_sub": g.row_a[:, 0].astype(np.int64),
        "row_a_dof": g.row_a[:, 1].astype(np.int64),
        "row_b_sub": np.concatenate([g.row_b[:, 0], -np.ones(g.nlam - g.n_glue)]).astype(np.int64),
        "weight": np.asarray(g.weight, dtype=np.float64),
    }
    out["l2g"], out["l2g_ptr"] = _concat([s.l2g[d] for d in range(m.nsub)])
    out["bf_rows"], out["bf_ptr"] = _concat([g.Bf[d]["rows"] for d in range(m.nsub)])
    out["bf_col"], _ = _concat([g.Bf[d]["col"] for d in range(m.nsub)])
    out["bf_sign"], _ = _concat([g.Bf[d]["val"].astype(np.int64) for d in range(m.nsub)])
    out["jcol"], out["jcol_ptr"] = _concat([g.Bf[d]["j_col_cmpr"] for d in range(m.nsub)])
    return out


def native_arrays(kw):
    import hfls_amd
    p = hfls_amd.Problem(**kw)
    a, b, dof, w = p.rows()
    out = {
        "dirichlet": np.asarray(p.dirichlet(), dtype=np.int64),
        "corners": np.asarray(p.corners(), dtype=np.int64),
        "row_a_sub": np.asarray(a, dtype=np.int64),
        "row_a_dof": np.asarray(dof, dtype=np.int64),
        "row_b_sub": np.asarray(b, dtype=np.int64),
        "weight": np.asarray(w, dtype=np.float64),
    }
    bf = [p.sub_bf(d) for d in range(p.nsub)]
    out["l2g"], out["l2g_ptr"] = _concat([p.sub_l2g(d) for d in range(p.nsub)])
    out["bf_rows"], out["bf_ptr"] = _concat([x[0] for x in bf])
    out["bf_col"], _ = _concat([x[1] for x in bf])
    out["bf_sign"], _ = _concat([x[2] for x in bf])
    out["jcol"], out["jcol_ptr"] = _concat([p.sub_bdofs(d) for d in range(p.nsub)])
    return out


def digests(arrays):
    """SHA-256 of each array's little-endian bytes (int64 or float64)."""
    out = {}
    for k in KEYS:
        v = arrays[k]
        v = v.astype("<f8") if v.dtype.kind == "f" else v.astype("<i8")
        out[k] = hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest()
    return out


def write_all():
    a0 = oracle_arrays(CUBE8_5E3)
    np.savez_compressed(os.path.join(HERE, "cube8_5e3_integer.npz"), **a0)
    a1 = oracle_arrays(CUBE512_10E3)
    meta = {"options": CUBE512_10E3, "nlam": int(len(a1["weight"])), "nsub": int(len(a1["l2g_ptr"]) - 1),
            "sha256": digests(a1)}
    with open(os.path.join(HERE, "cube512_10e3_integer_sha256.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    from oracle.mesh import opts
    from oracle.feti import HTFETI
    h = HTFETI(opts(eps_iter="1e-10", **CUBE8_10E3), sparse=True)
    r = h.pcpg(eps=1e-10, max_iter=1000)
    r8 = h.pcpg(eps=1e-8, max_iter=1000)
    np.savez(os.path.join(HERE, "cube8_10e3_u.npz"), u=np.asarray(r["u"], dtype=np.float64),
             iterations_1e8=np.int64(r8["iterations"]), iterations_1e10=np.int64(r["iterations"]))


if __name__ == "__main__":
    write_all()
    print("fixtures written to", HERE)
