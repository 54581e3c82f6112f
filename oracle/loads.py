"""Seeded random surface tractions on the x = 1 face (oracle; test infrastructure).

Multi-load-case workloads need loads other than the reference's single body force
(Data.cpp:315-318).  Definition (the product's traction_loads follows it value for
value, so the loads agree bit for bit: the same splitmix64 integers, the same
sequence of IEEE operations, and glibc's log/cos behind both math.log and
std::log):

  * t(c, jj, kk) ~ N(0,1) per node of the x = 1 face and component c, Box-Muller
    over a splitmix64 stream seeded with seed ^ (0xD1B54A32D192ED03 * (case + 1)),
    drawn kk outer, jj, then c (two uniforms per normal);
  * three Jacobi passes of the 5-point face smoother
        t(p) <- (t(p) + t(jj-1) + t(jj+1) + t(kk-1) + t(kk+1)) / (1 + #neighbours);
  * lumped nodal forces per subdomain: f = t * (cnt * hy * hz * 0.25), cnt the
    number of the subdomain's face elements at the node, hy = 1/NY, hz = 1/NZ.

Pinned by: the total force over all subdomains equals the face quadrature
sum_nodes t * (cnt_global * hy * hz / 4); smoothing preserves a constant field;
the raw normals have mean ~0 and variance ~1 (tests/test_oracle_loads.py).
"""
import math

import numpy as np

from .mesh import MASK64, splitmix64


def face_field(NE, seed, case):
    """Smoothed traction t[c, kk, jj] on the x = 1 face for one load case."""
    NX, NY, NZ = NE
    ny1, nz1 = NY + 1, NZ + 1
    st = (seed ^ ((0xD1B54A32D192ED03 * (case + 1)) & MASK64)) & MASK64
    t = np.empty((3, nz1, ny1))
    for kk in range(nz1):
        for jj in range(ny1):
            for c in range(3):
                st, z0 = splitmix64(st)
                st, z1 = splitmix64(st)
                u0 = ((z0 >> 11) + 0.5) / 9007199254740992.0
                u1 = ((z1 >> 11) + 0.5) / 9007199254740992.0
                t[c, kk, jj] = math.sqrt(-2.0 * math.log(u0)) * math.cos(2.0 * math.pi * u1)
    return smooth(t, 3)


def smooth(t, passes):
    """Jacobi passes of the 5-point face smoother on t[c, kk, jj]."""
    nz1, ny1 = t.shape[1], t.shape[2]
    m = np.ones((nz1, ny1))
    m[:, 1:] += 1.0
    m[:, :-1] += 1.0
    m[1:, :] += 1.0
    m[:-1, :] += 1.0
    for _ in range(passes):
        s = t.copy()
        s[:, :, 1:] += t[:, :, :-1]
        s[:, :, :-1] += t[:, :, 1:]
        s[:, 1:, :] += t[:, :-1, :]
        s[:, :-1, :] += t[:, 1:, :]
        t = s / m
    return t


def traction_loads(mesh, subs, seed, cases):
    """Packed local loads (list over cases of concatenated per-subdomain vectors,
    l2g order) for the given load-case indices."""
    NX, NY, NZ = mesh.NE
    hy, hz = 1.0 / NY, 1.0 / NZ
    nx1, nxy1 = NX + 1, (NX + 1) * (NY + 1)
    out = []
    for case in cases:
        t = face_field(mesh.NE, seed, case)
        parts = []
        for d in range(mesh.nsub):
            l2g = subs.l2g[d]
            f = np.zeros(len(l2g))
            el = subs.elems[d]
            ei = el % NX
            face = el[ei == NX - 1]
            if len(face):
                ej = (face // NX) % NY
                ek = face // (NX * NY)
                cnt = {}
                for a, b in zip(ej, ek):
                    for jj, kk in ((a, b), (a + 1, b), (a + 1, b + 1), (a, b + 1)):
                        cnt[(int(jj), int(kk))] = cnt.get((int(jj), int(kk)), 0) + 1
                for (jj, kk), c_ in cnt.items():
                    w = float(c_) * hy * hz * 0.25
                    node = NX + jj * nx1 + kk * nxy1
                    for c in range(3):
                        loc = np.searchsorted(l2g, 3 * node + c)
                        f[loc] = t[c, kk, jj] * w
            parts.append(f)
        out.append(np.concatenate(parts))
    return out
