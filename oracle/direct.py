"""Global FP64 direct solve (brute-force pin for the FETI chain).  TEST INFRASTRUCTURE ONLY.

Assembles the undecomposed stiffness of the whole cube from the same element
routine (oracle.fe.hex8), removes the Dirichlet DOFs and solves with SuperLU.
The displacement of a correct FETI solve equals this one to the PCPG tolerance.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from .fe import hex8, load_vector


def global_solve(mesh, f_sub=None, subs=None):
    """Return u (3*nPoints) solving K u = f with u = 0 on the Dirichlet DOFs.

    If f_sub (per-subdomain load vectors in l2g order) is given, the global load
    is their sum; otherwise the body force of the options is assembled."""
    o = mesh.o
    mu = float(o["poissons_ratio"])
    E_el = mesh.young_per_element()
    nodes = mesh.elements
    ieq = np.concatenate([3 * nodes, 3 * nodes + 1, 3 * nodes + 2], axis=1)
    Ke, fe = hex8(mesh.points[nodes], E_el, mu, load_vector(o))
    n = 3 * mesh.nPoints
    rows = np.repeat(ieq, 24, axis=1).ravel()
    cols = np.tile(ieq, (1, 24)).ravel()
    K = sp.coo_matrix((Ke.ravel(), (rows, cols)), shape=(n, n)).tocsr()
    f = np.zeros(n)
    if f_sub is None:
        np.add.at(f, ieq.ravel(), fe.ravel())
    else:
        for d in range(len(f_sub)):
            np.add.at(f, subs.l2g[d], f_sub[d])
    free = np.ones(n, dtype=bool)
    free[mesh.dirichlet_dofs] = False
    idx = np.nonzero(free)[0]
    Kff = K[idx][:, idx].tocsc()
    u = np.zeros(n)
    u[idx] = spla.spsolve(Kff, f[idx])
    return u, K, f
