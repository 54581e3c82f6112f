"""Pins of the oracle's seeded face tractions (oracle/loads.py) by properties the
definition fixes: the smoother keeps a constant field, the lumped nodal forces sum
to the face quadrature of the field, only x = 1 face DOFs are loaded, different
cases and seeds differ, and the raw Box-Muller normals are standard normal."""
import math

import numpy as np

from oracle.mesh import Mesh, opts, splitmix64, MASK64
from oracle.fe import Subdomains
from oracle.loads import face_field, smooth, traction_loads

KW = dict(Cx=2, Cy=1, Cz=1, Nx=2, Ny=2, Nz=2, nx=3, ny=3, nz=4, bc="clamp_x0")


def test_smoother_keeps_constants_and_averages():
    t = np.full((3, 5, 4), 2.5)
    assert np.array_equal(smooth(t, 3), t)
    t = np.zeros((1, 3, 3))
    t[0, 1, 1] = 5.0
    s = smooth(t, 1)
    assert s[0, 1, 1] == 1.0 and s[0, 0, 1] == 5.0 / 4.0 and s[0, 0, 0] == 0.0


def test_total_force_is_face_quadrature():
    m = Mesh(opts(**KW))
    s = Subdomains(m)
    NX, NY, NZ = m.NE
    f = traction_loads(m, s, 11, [0])[0]
    t = face_field(m.NE, 11, 0)
    w = np.ones((NZ + 1, NY + 1))
    w[:, 0] = w[:, -1] = 0.5
    w[0, :] *= 0.5
    w[-1, :] *= 0.5
    total = np.zeros(3)
    off = 0
    for d in range(m.nsub):
        l2g = s.l2g[d]
        fd = f[off:off + len(l2g)]
        off += len(l2g)
        for c in range(3):
            total[c] += fd[l2g % 3 == c].sum()
        nodes = l2g[fd != 0] // 3
        assert np.all(m.grid[nodes, 0] == NX)  # only the x = 1 face is loaded
    expect = np.array([(t[c] * w).sum() / (NY * NZ) for c in range(3)])
    assert np.allclose(total, expect, rtol=1e-12, atol=1e-15)


def test_cases_and_seeds_differ():
    NE = (4, 3, 5)
    a, b, c = face_field(NE, 1, 0), face_field(NE, 1, 1), face_field(NE, 2, 0)
    assert not np.array_equal(a, b) and not np.array_equal(a, c)
    assert np.array_equal(a, face_field(NE, 1, 0))


def test_box_muller_normals_are_standard():
    st = (5 ^ ((0xD1B54A32D192ED03 * 1) & MASK64)) & MASK64
    z = []
    for _ in range(20000):
        st, a = splitmix64(st)
        st, b = splitmix64(st)
        u0 = ((a >> 11) + 0.5) / 9007199254740992.0
        u1 = ((b >> 11) + 0.5) / 9007199254740992.0
        z.append(math.sqrt(-2.0 * math.log(u0)) * math.cos(2.0 * math.pi * u1))
    z = np.array(z)
    assert abs(z.mean()) < 0.03 and abs(z.var() - 1.0) < 0.04
