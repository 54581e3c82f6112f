"""Hex8 (ANSYS solid45) element and subdomain assembly.  TEST INFRASTRUCTURE ONLY.

Restates Data::stf_mtrx_solid45 (Data.cpp:297-490), inverse_matrix_3x3
(Data.cpp:276-294), fe_assemb_local_K_f (Data.cpp:25-54), buildMappingStruct
(Data.cpp:111-144), feti_symbolic (Data.cpp:147-228) and feti_numeric
(Data.cpp:231-273).

Parity: pinned by closed forms in tests/test_oracle_fe.py -- symmetry, the six
rigid-body modes in the null space, h-scaling, the exact strain energy of affine
fields (2x2x2 Gauss is exact on a brick), sum(f_e) = rho*a*V, and the closed-form
upper-CSR nnz (9(3m-2)^3 + n)/2.
"""
import numpy as np
import scipy.sparse as sp

PT = 0.577350269189626  # Data.cpp:349 (as printed there)
# gauss points (Data.cpp:372-396): r = gp[i], s = gp[i+8], t = gp[i+16]
_GP = np.array([
    -PT, -PT, -PT, -PT, PT, PT, PT, PT,
    -PT, -PT, PT, PT, -PT, -PT, PT, PT,
    -PT, PT, -PT, PT, -PT, PT, -PT, PT])


def _shape(r, s, t):
    dNr = 0.125 * np.array([-(1 - s) * (1 - t), (1 - s) * (1 - t), (1 + s) * (1 - t), -(1 + s) * (1 - t),
                            -(1 - s) * (1 + t), (1 - s) * (1 + t), (1 + s) * (1 + t), -(1 + s) * (1 + t)])
    dNs = 0.125 * np.array([-(1 - r) * (1 - t), -(1 + r) * (1 - t), (1 + r) * (1 - t), (1 - r) * (1 - t),
                            -(1 - r) * (1 + t), -(1 + r) * (1 + t), (1 + r) * (1 + t), (1 - r) * (1 + t)])
    dNt = 0.125 * np.array([-(1 - r) * (1 - s), -(1 + r) * (1 - s), -(1 + r) * (1 + s), -(1 - r) * (1 + s),
                            (1 - r) * (1 - s), (1 + r) * (1 - s), (1 + r) * (1 + s), (1 - r) * (1 + s)])
    N = 0.125 * np.array([(1 - r) * (1 - s) * (1 - t), (r + 1) * (1 - s) * (1 - t), (r + 1) * (s + 1) * (1 - t),
                          (1 - r) * (s + 1) * (1 - t), (1 - r) * (1 - s) * (t + 1), (r + 1) * (1 - s) * (t + 1),
                          (r + 1) * (s + 1) * (t + 1), (1 - r) * (s + 1) * (t + 1)])
    return dNr, dNs, dNt, N


def elasticity_C(E, mu):
    """6x6 Voigt matrix, order xx,yy,zz,xy,yz,xz (Data.cpp:341-361)."""
    c = E / ((1. + mu) * (1. - 2. * mu))
    C = np.zeros((6, 6))
    C[0, 0] = C[1, 1] = C[2, 2] = (1. - mu) * c
    C[0, 1] = C[0, 2] = C[1, 0] = C[1, 2] = C[2, 0] = C[2, 1] = mu * c
    C[3, 3] = C[4, 4] = C[5, 5] = (0.5 - mu) * c
    return C


def hex8(coords, E, mu, acc):
    """Element stiffness and load for a batch of elements.

    coords: (ne, 8, 3); E: (ne,); acc: 3-vector rho*a.  Returns K (ne,24,24) with
    local DOF order [x0..x7, y0..y7, z0..z7] (Data.cpp:42-44) and f (ne,24)."""
    coords = np.asarray(coords, dtype=np.float64)
    ne = coords.shape[0]
    K = np.zeros((ne, 24, 24))
    f = np.zeros((ne, 24))
    C1 = elasticity_C(1.0, mu)
    for g in range(8):
        r, s, t = _GP[g], _GP[g + 8], _GP[g + 16]
        dNr, dNs, dNt, N = _shape(r, s, t)
        x, y, z = coords[:, :, 0], coords[:, :, 1], coords[:, :, 2]
        A = np.stack([x @ dNr, x @ dNs, x @ dNt, y @ dNr, y @ dNs, y @ dNt, z @ dNr, z @ dNs, z @ dNt], axis=1)
        # inverse_matrix_3x3 (Data.cpp:276-294), element-wise
        det = (A[:, 0] * (A[:, 4] * A[:, 8] - A[:, 5] * A[:, 7])
               - A[:, 3] * (A[:, 1] * A[:, 8] - A[:, 7] * A[:, 2])
               + A[:, 6] * (A[:, 1] * A[:, 5] - A[:, 4] * A[:, 2]))
        inv = 1.0 / det
        iA = np.stack([
            (A[:, 4] * A[:, 8] - A[:, 5] * A[:, 7]) * inv,
            -(A[:, 3] * A[:, 8] - A[:, 6] * A[:, 5]) * inv,
            (A[:, 3] * A[:, 7] - A[:, 6] * A[:, 4]) * inv,
            -(A[:, 1] * A[:, 8] - A[:, 7] * A[:, 2]) * inv,
            (A[:, 0] * A[:, 8] - A[:, 6] * A[:, 2]) * inv,
            -(A[:, 0] * A[:, 7] - A[:, 1] * A[:, 6]) * inv,
            (A[:, 1] * A[:, 5] - A[:, 2] * A[:, 4]) * inv,
            -(A[:, 0] * A[:, 5] - A[:, 2] * A[:, 3]) * inv,
            (A[:, 0] * A[:, 4] - A[:, 1] * A[:, 3]) * inv], axis=1)
        dNx = iA[:, 0:1] * dNr + iA[:, 3:4] * dNs + iA[:, 6:7] * dNt
        dNy = iA[:, 1:2] * dNr + iA[:, 4:5] * dNs + iA[:, 7:8] * dNt
        dNz = iA[:, 2:3] * dNr + iA[:, 5:6] * dNs + iA[:, 8:9] * dNt
        # strain-displacement layout Gt (Data.cpp:451-464)
        Gt = np.zeros((ne, 6, 24))
        Gt[:, 0, 0:8] = dNx
        Gt[:, 1, 8:16] = dNy
        Gt[:, 2, 16:24] = dNz
        Gt[:, 3, 0:8] = dNy
        Gt[:, 3, 8:16] = dNx
        Gt[:, 4, 8:16] = dNz
        Gt[:, 4, 16:24] = dNy
        Gt[:, 5, 0:8] = dNz
        Gt[:, 5, 16:24] = dNx
        CG = np.einsum("kl,elj->ekj", C1, Gt)
        K += np.einsum("eki,ekj->eij", Gt, CG) * (det * E)[:, None, None]
        for c in range(3):
            f[:, 8 * c:8 * c + 8] += acc[c] * N[None, :] * det[:, None]
    return K, f


def load_vector(o):
    """rho*a: reference (Data.cpp:315-318) is density 2, a=(0,0,98.1);
    "body" is the per-unit-volume force (0,0,body_force)."""
    if o["load"] == "reference":
        return np.array([0.0, 0.0, 2.0 * 98.10])
    if o["load"] in ("body", "traction"):
        return np.array([0.0, 0.0, float(o["body_force"])])
    raise ValueError(o["load"])


class Subdomains:
    """Per-subdomain K_d (full symmetric scipy CSR), f_d, l2g_d (Data.cpp:25-273)."""

    def __init__(self, mesh):
        o = mesh.o
        mu = float(o["poissons_ratio"])
        Esub = mesh.young_per_sub()
        Eel = mesh.young_per_element()
        acc = load_vector(o)
        self.mesh = mesh
        self.E = Esub
        self.l2g, self.K, self.f, self.elems = [], [], [], []
        for d in range(mesh.nsub):
            el = np.nonzero(mesh.part == d)[0]  # ascending global element index
            nodes = mesh.elements[el]  # (ne, 8)
            ieq = np.concatenate([3 * nodes, 3 * nodes + 1, 3 * nodes + 2], axis=1)  # (ne, 24)
            l2g = np.unique(ieq)
            loc = np.searchsorted(l2g, ieq)
            Ke, fe = hex8(mesh.points[nodes], Eel[el], mu, acc)
            n = len(l2g)
            rows = np.repeat(loc, 24, axis=1).ravel()
            cols = np.tile(loc, (1, 24)).ravel()
            Kd = sp.coo_matrix((Ke.ravel(), (rows, cols)), shape=(n, n)).tocsr()
            Kd.sum_duplicates()
            fd = np.zeros(n)
            np.add.at(fd, loc.ravel(), fe.ravel())
            self.l2g.append(l2g)
            self.K.append(Kd)
            self.f.append(fd)
            self.elems.append(el)

    @staticmethod
    def upper_csr_pattern(Kd):
        """Upper-triangular CSR pattern with sorted columns (Data.cpp:168-225)."""
        U = sp.triu(Kd, format="csr")
        U.sort_indices()
        return U.indptr.copy(), U.indices.copy()
