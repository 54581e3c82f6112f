"""Kernel detection and regularisation.  TEST INFRASTRUCTURE ONLY.

Restates Matrix::get_kernel_from_K (Matrix.cpp:1982-2448), GramSchmidtOrtho
(Matrix.cpp:2450-2472), getNullPivots (Matrix.cpp:835-884) and the diagonal
regularisation of Matrix::num_factor(R) (Matrix.cpp:1029-1105).

Two s-set rules:
  * "reference": a permutation of all DOFs, last ``sc_size`` (=50) form the s-set
    (Matrix.cpp:2054,2163-2183).  The reference seeds it with time(NULL); here a
    seeded numpy permutation is used, so runs are reproducible.
  * "geometric": the deterministic rule the product uses -- the 8 B-touched nodes
    extreme along (+-1,+-1,+-1) in integer grid coordinates (ties -> lowest local
    node), all their B-touched DOFs.  Fixing DOFs are then chosen inside the s-set
    by pivoted Cholesky of R_s R_s^T (basis independent; ties within 1e-9
    relative go to the lowest index), so that the regularised DOFs lie on the
    interface and K_reg^-1 restricted to it is (S_bb + D)^-1.

Parity: defect == 6, ||K R||/(||K|| ||R||) <= 1e-10, R^T R = I and the subspace
angle to the analytic rigid-body modes (Data.cpp:61-109) are checked in
tests/test_oracle_kernel.py; ||K - K K_reg^-1 K|| / ||K|| <= 1e-10 pins the
regularisation (Appendix fact: K_reg^-1 is a generalised inverse).
"""
import numpy as np
import scipy.linalg as sla

JUMP = 1.0e-5   # Matrix.cpp:2065
TWENTY = 20     # Matrix.cpp:2059
SC_SIZE = 50    # Matrix.cpp:2054


def gram_schmidt(R):
    """Classical Gram-Schmidt, column by column (Matrix.cpp:2450-2472)."""
    R = R.copy()
    for j in range(R.shape[1]):
        w = R[:, j].copy()
        for i in range(j):
            w -= R[:, i] * np.dot(w, R[:, i])
        R[:, j] = w / np.sqrt(np.dot(w, w))
    return R


def null_pivots(R):
    """Matrix::getNullPivots (Matrix.cpp:835-884): full-pivot elimination on R,
    pivots swapped to the bottom; returns the sorted pivot rows."""
    N = R.copy()
    n_row, n_col = N.shape
    piv = np.arange(n_row)
    for j in range(n_col):
        sub = np.abs(N[:n_row, :n_col - j])  # columns 0..n_col-1-j (max_element over the prefix)
        # max_element over column-major storage with <= comparison returns the
        # LAST maximum in storage order (compareDouble is fabs(i) <= fabs(j)).
        flat = sub.ravel(order="F")
        m = flat.max()
        I = int(np.nonzero(flat == m)[0][-1])
        colInd, rowInd = I // n_row, I % n_row
        N[[n_row - 1 - j, rowInd], :n_col - j] = N[[rowInd, n_row - 1 - j], :n_col - j]
        piv[[rowInd, n_row - 1 - j]] = piv[[n_row - 1 - j, rowInd]]
        if n_col - 1 - j != colInd:
            N[:, [n_col - 1 - j, colInd]] = N[:, [colInd, n_col - 1 - j]]
        pivot = N[n_row - 1 - j, n_col - 1 - j]
        for J in range(n_col - j - 1):
            N[:n_row - j, J] -= N[:n_row - j, n_col - 1 - j] * N[n_row - 1 - j, J] / pivot
    return np.sort(piv[n_row - n_col:][::-1])


def projector_pivots(Rs, ndef=6, tie=1e-9):
    """Basis-independent fixing rows: pivoted Cholesky of P = Rs Rs^T (rank ndef)."""
    P = Rs @ Rs.T
    n = P.shape[0]
    d = np.diag(P).copy()
    L = np.zeros((n, ndef))
    sel = []
    for k in range(ndef):
        cand = [i for i in range(n) if i not in sel]
        m = max(d[i] for i in cand)
        p = min(i for i in cand if d[i] >= m * (1.0 - tie))
        sel.append(p)
        L[:, k] = (P[:, p] - L[:, :k] @ L[p, :k]) / np.sqrt(d[p])
        d -= L[:, k] ** 2
    return np.array(sorted(sel), dtype=np.int64)


def sset_geometric(grid_nodes, bdofs):
    """Deterministic s-set: for each direction in (+-1,+-1,+-1) (x fastest, + before
    -), the node with a B-touched DOF maximising dir . (i,j,k); ties -> lowest
    local node.  Returns sorted local DOFs: every B-touched DOF of those nodes.

    grid_nodes: (n_nodes_local, 3) integer grid coordinates of local nodes (local
    node = local DOF // 3); bdofs: sorted B-touched local DOFs."""
    bnodes = np.unique(bdofs // 3)
    g = grid_nodes[bnodes]
    chosen = []
    for sz in (1, -1):
        for sy in (1, -1):
            for sx in (1, -1):
                v = sx * g[:, 0] + sy * g[:, 1] + sz * g[:, 2]
                m = v.max()
                chosen.append(int(bnodes[np.nonzero(v == m)[0][0]]))
    chosen = set(chosen)
    bset = set(int(x) for x in bdofs)
    s = sorted(3 * n + c for n in chosen for c in range(3) if 3 * n + c in bset)
    return np.array(s, dtype=np.int64)


def kernel_from_K(Kd, sset=None, seed=0, dense=True):
    """get_kernel_from_K (Matrix.cpp:1982-2448).

    Kd: symmetric (scipy sparse or dense) FP64.  sset: explicit s-set (sorted
    local DOFs) or None for the reference's random s-set of 50.  Returns (R,
    defect, sset) with R orthonormalised (GramSchmidtOrtho)."""
    K = Kd.toarray() if hasattr(Kd, "toarray") else np.asarray(Kd)
    n = K.shape[0]
    # diagonal scaling K_modif[i,j] = K[i,j]/sqrt(K[i,i]K[j,j]) (Matrix.cpp:2137-2149)
    dg = np.sqrt(np.diag(K).copy())
    Ks = K / dg[:, None] / dg[None, :]
    if sset is None:
        sc = SC_SIZE if n >= SC_SIZE else n
        perm = np.random.default_rng(seed).permutation(n)
        r = np.sort(perm[:n - sc])
        s = np.sort(perm[n - sc:])
    else:
        s = np.asarray(sset)
        r = np.setdiff1d(np.arange(n), s)
    Krr = Ks[np.ix_(r, r)]
    Krs = Ks[np.ix_(r, s)]
    S = Ks[np.ix_(s, s)].copy()
    if len(r):
        c = sla.cho_factor(Krr, lower=True)
        S -= Krs.T @ sla.cho_solve(c, Krs)
    S = 0.5 * (S + S.T)
    W, Z = np.linalg.eigh(S)
    itmax = min(TWENTY, len(s))
    defect = 0
    for i in range(itmax - 1, 0, -1):
        if abs(W[i - 1] / W[i]) < JUMP:
            defect = i
            break
    Rs = Z[:, :defect]
    R = np.zeros((n, defect))
    if len(r):
        R[r] = -sla.cho_solve(c, Krs @ Rs)
    R[s] = Rs
    R /= dg[:, None]
    return gram_schmidt(R), defect, s


def regularize(Kd, fix):
    """K_reg = K with K[p,p] doubled at the fixing DOFs (Matrix.cpp:1033-1046)."""
    K = Kd.toarray() if hasattr(Kd, "toarray") else np.array(Kd, copy=True)
    K = np.array(K, copy=True)
    K[fix, fix] *= 2.0
    return K
