"""Full-size displacement fixtures written from the oracle (sparse LU K_reg, implicit K+).

Run from the repository root, one workload per call (each takes minutes to tens
of minutes on the CPU):

    python tests/golden/golden_full.py cube512_10e3     # 8 clusters x 4^3 subdomains of 10^3 hex8
    python tests/golden/golden_full.py cube64_12e3      # 1 cluster of 4^3 subdomains of 12^3 hex8
    python tests/golden/golden_full.py uneven512        # 8 clusters x 4^3, slabs from {6,8,10,12,14}

Each writes ``<name>_u.npz`` next to this file with, per load case:

* ``u_sample`` -- the oracle's displacement at a seeded sample of global DOFs
  (``sample``, 16,384 sorted DOFs drawn without replacement) at eps 1e-10;
* ``u_norm`` (|u|_2 over all DOFs), ``u_sum`` (sum of u per component x, y, z);
* ``iterations_1e8`` / ``iterations_1e10`` -- F applications of the PCPG loop
  (Cluster.cpp:1422-1450) until sqrt(g'Pz) < eps sqrt(g0'Pz0) at 1e-8 and 1e-10,
  both from the one run at 1e-10 (the iteration does not depend on eps before
  it stops, so the first history entry below 1e-8 is the count at 1e-8).

Load cases: the body force of the options (case name "body") and seeded face
tractions (``oracle/loads.py``; seed 20240611, case indices as listed), the
same loads the product generates bit for bit.
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

SEED = 20240611
NSAMPLE = 16384
BASE = dict(bc="clamp_x0", young_modulus="2.1e5", poissons_ratio="0.3", load="body", body_force="-1.0",
            typeBc="ker", preconditioner="Dirichlet_explicit")
WORKLOADS = {
    # the 512-subdomain cube: 2x2x2 clusters x 4x4x4 subdomains x 10^3 elements
    "cube512_10e3": (dict(BASE, Cx=2, Cy=2, Cz=2, Nx=4, Ny=4, Nz=4, nx=10, ny=10, nz=10, material="uniform"),
                     ["body", 0, 1]),
    # one cluster of the 512 x 12^3 cube's shape: 4x4x4 subdomains x 12^3 elements, four face tractions
    "cube64_12e3": (dict(BASE, Nx=4, Ny=4, Nz=4, nx=12, ny=12, nz=12, material="uniform"), [0, 1, 2, 3]),
    # uneven batching: per-slab element counts drawn from {6,8,10,12,14}, lognormal E per subdomain
    # (sigma 1.5, clipped to a 1e3 contrast)
    "uneven512": (dict(BASE, Cx=2, Cy=2, Cz=2, Nx=4, Ny=4, Nz=4, slab_choices="6,8,10,12,14", seed=str(SEED),
                       material="lognormal", mat_sigma="1.5", mat_contrast="1e3"), ["body"]),
}


def split(subs, fpacked):
    out, off = [], 0
    for l2g in subs.l2g:
        out.append(np.asarray(fpacked[off:off + len(l2g)]))
        off += len(l2g)
    return out


def first_below(hist, eps):
    for i, r in enumerate(hist):
        if r < eps:
            return i
    return -1


def write(name):
    from oracle.mesh import opts
    from oracle.feti import HTFETI
    from oracle.loads import traction_loads
    kw, cases = WORKLOADS[name]
    t0 = time.time()
    h = HTFETI(opts(eps_iter="1e-10", **kw), sparse=True)
    print("%s: set-up %.0f s, %d subdomains, %d dual rows" % (name, time.time() - t0, h.nsub, h.nlam), flush=True)
    n = 3 * h.mesh.nPoints
    sample = np.sort(np.random.default_rng(SEED).choice(n, size=min(NSAMPLE, n), replace=False)).astype(np.int64)
    out = {"sample": sample, "options": np.array(json.dumps(kw, sort_keys=True)),
           "cases": np.array(json.dumps([str(c) for c in cases])), "traction_seed": np.int64(SEED)}
    us, norms, sums, it8, it10 = [], [], [], [], []
    for c in cases:
        f = None if c == "body" else split(h.subs, traction_loads(h.mesh, h.subs, SEED, [c])[0])
        t1 = time.time()
        r = h.pcpg(f=f, eps=1e-10, max_iter=2000, history=True)
        u = np.asarray(r["u"], dtype=np.float64)
        us.append(u[sample])
        norms.append(np.linalg.norm(u))
        sums.append([u[0::3].sum(), u[1::3].sum(), u[2::3].sum()])
        it8.append(first_below(r["history"], 1e-8))
        it10.append(r["iterations"])
        print("%s case %s: %d iterations to 1e-10 (%d to 1e-8), |u| %.6e, %.0f s" %
              (name, c, it10[-1], it8[-1], norms[-1], time.time() - t1), flush=True)
    out.update(u_sample=np.array(us), u_norm=np.array(norms), u_sum=np.array(sums),
               iterations_1e8=np.array(it8, dtype=np.int64), iterations_1e10=np.array(it10, dtype=np.int64))
    np.savez_compressed(os.path.join(HERE, name + "_u.npz"), **out)
    print("%s: written, %.0f s in all" % (name, time.time() - t0), flush=True)


if __name__ == "__main__":
    for nm in sys.argv[1:]:
        write(nm)
