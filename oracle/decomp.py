"""Interfaces, gluing matrices and weights.  TEST INFRASTRUCTURE ONLY.

Restates Cluster::create_cluster_constraints (Cluster.cpp:682-769),
create_Bc_or_Bf_in_CSR (Cluster.cpp:824-928), create_Bc_weightedAverages_in_COO
(Cluster.cpp:772-821), matrix_Bx_COO2CSR (Cluster.cpp:931-959) and
Matrix::test_of_Bc_constraints (Matrix.cpp:1538-1571).

The reference finds neighbours by pairwise set intersection of the subdomain DOF
sets; here the same sets are formed by grouping (DOF, subdomain) pairs, which
yields identical ascending lists.  Clusters (an extension) restrict Bc to pairs
inside one cluster; Bf is global over all clusters.

Extension: typeBc="cor+edge" -- the `cor` rows plus, per cluster-interior edge of
the subdomain grid, one averaged row (+-1/m over the m interior edge nodes) per
component between the first subdomain holding the edge and each other one (every
pair without Bc_fullRank).  Pinned by dim ker(Gc^T Gc) = 6, the displacement of
`ker`, and a hand count of the edge rows (tests/test_options.py).

Parity: row order, columns, signs, weights and compressed-row maps are pinned by
closed-form counts (sum over DOFs of m(m-1)/2 plus Dirichlet rows), the per-row
structure (+1/-1 pairs or a single +1), hand-enumerated 2-subdomain cases in
tests/test_oracle_decomp.py, and the committed fixtures of tests/golden.
"""
import numpy as np


class Gluing:
    def __init__(self, mesh, subs):
        self.mesh = mesh
        nsub = mesh.nsub
        l2g = subs.l2g
        # (dof, sub) pairs sorted by dof then sub -> ascending subdomain lists
        g = np.concatenate(l2g)
        d = np.concatenate([np.full(len(l2g[i]), i, dtype=np.int64) for i in range(nsub)])
        order = np.lexsort((d, g))
        g, d = g[order], d[order]
        self.pair_g, self.pair_d = g, d
        ug, start, mult = np.unique(g, return_index=True, return_counts=True)
        self.mult = np.zeros(3 * mesh.nPoints, dtype=np.int64)
        self.mult[ug] = mult
        self.first = np.zeros(3 * mesh.nPoints, dtype=np.int64)
        self.first[ug] = start
        dir_set = np.zeros(3 * mesh.nPoints, dtype=bool)
        dir_set[mesh.dirichlet_dofs] = True
        self.is_dir = dir_set

        # ---- Bf (Cluster.cpp:846-924, addDirConstr=true, full_rank=false) ----
        # triples (d, g, s) with s > d sharing a non-Dirichlet DOF g, ordered by
        # d, then ascending g (map order of interface[d]), then ascending s.
        rows_d, rows_g, rows_s = [], [], []
        shared = ug[mult > 1]
        for gg in shared:
            if dir_set[gg]:
                continue
            s0 = self.first[gg]
            ss = d[s0:s0 + self.mult[gg]]
            for a in range(len(ss)):
                for b in range(a + 1, len(ss)):
                    rows_d.append(ss[a]); rows_g.append(gg); rows_s.append(ss[b])
        rows_d = np.array(rows_d, dtype=np.int64)
        rows_g = np.array(rows_g, dtype=np.int64)
        rows_s = np.array(rows_s, dtype=np.int64)
        o = np.lexsort((rows_s, rows_g, rows_d))
        rows_d, rows_g, rows_s = rows_d[o], rows_g[o], rows_s[o]
        n_glue = len(rows_d)
        w_glue = 1.0 / (self.mult[rows_g].astype(np.float64))  # 1/(|interface[d][g]|+1)
        # Dirichlet rows: for d, for DOF in DirichletDOFs order, if in l2g[d]
        dd, dg = [], []
        for sdom in range(nsub):
            mask = np.isin(mesh.dirichlet_dofs, l2g[sdom])
            for gg in mesh.dirichlet_dofs[mask]:
                dd.append(sdom); dg.append(gg)
        dd = np.array(dd, dtype=np.int64)
        dg = np.array(dg, dtype=np.int64)
        self.n_glue = n_glue
        self.nlam = n_glue + len(dd)
        self.weight = np.concatenate([w_glue, np.ones(len(dd))])
        # row entries: (sub, global dof, sign)
        self.row_a = np.stack([np.concatenate([rows_d, dd]), np.concatenate([rows_g, dg])], axis=1)
        self.row_b = np.stack([rows_s, rows_g], axis=1)  # only for the first n_glue rows

        # per-subdomain compressed Bf (Cluster.cpp:931-959): sorted global rows,
        # local column and value per compressed row
        self.Bf = []
        for sdom in range(nsub):
            ra = np.nonzero(self.row_a[:, 0] == sdom)[0]
            rb = np.nonzero(self.row_b[:, 0] == sdom)[0]
            rows = np.concatenate([ra, rb])
            gdof = np.concatenate([self.row_a[ra, 1], self.row_b[rb, 1]])
            val = np.concatenate([np.ones(len(ra)), -np.ones(len(rb))])
            o2 = np.argsort(rows, kind="stable")
            rows, gdof, val = rows[o2], gdof[o2], val[o2]
            col = np.searchsorted(l2g[sdom], gdof)
            self.Bf.append({"rows": rows, "col": col, "val": val,
                            "j_col_cmpr": np.unique(col)})

    # ------------------------------------------------------------------
    def subs_of(self, gg):
        s0 = self.first[gg]
        return self.pair_d[s0:s0 + self.mult[gg]]

    def interfaces(self, i):
        """data.interfaces[i]: list of (j, sorted shared global DOFs) for j > i
        (Cluster.cpp:707-727)."""
        mesh = self.mesh
        out = {}
        l2g_i = self._l2g[i]
        shared = l2g_i[self.mult[l2g_i] > 1]
        for gg in shared:
            for j in self.subs_of(gg):
                if j > i:
                    out.setdefault(int(j), []).append(int(gg))
        return [(j, np.array(out[j], dtype=np.int64)) for j in sorted(out)]

    def attach_l2g(self, l2g):
        self._l2g = l2g


def bc_rank_full(A):
    """Matrix::test_of_Bc_constraints (Matrix.cpp:1538-1571): singular values of
    A A^T (descending); a ratio S[i+1]/S[i] < 1e-4 flags a defect."""
    M = A @ A.T
    S = np.linalg.svd(M, compute_uv=False)
    for i in range(len(S) - 1):
        if abs(S[i + 1] / S[i]) < 1e-4:
            return False
    return True


def build_Bc(mesh, subs, glue, R, typeBc="ker", full_rank=True):
    """Cluster-level gluing Bc, per cluster with cluster-local row numbering.

    Returns a list over clusters of dicts {"nlam": n, "rows": {d: (rows, cols, vals)}}
    where rows/cols/vals are COO in reference push order (row-major, ascending)."""
    l2g = subs.l2g
    glue.attach_l2g(l2g)
    out = []
    for cl in range(mesh.ncl):
        d0 = cl * mesh.nsub_cl
        d1 = d0 + mesh.nsub_cl
        ent = {d: ([], [], []) for d in range(d0, d1)}
        cnt = 0
        if typeBc == "ker":
            # Cluster.cpp:779-819
            for i in range(d0, d1):
                for (j, dofs) in glue.interfaces(i):
                    if j >= d1:
                        continue
                    li = np.searchsorted(l2g[i], dofs)
                    lj = np.searchsorted(l2g[j], dofs)
                    A = R[i][li, :].T  # (6, n_com)
                    if full_rank and not bc_rank_full(A):
                        continue
                    for k in range(A.shape[0]):
                        ent[i][0].extend([cnt] * len(dofs)); ent[i][1].extend(li); ent[i][2].extend(A[k])
                        ent[j][0].extend([cnt] * len(dofs)); ent[j][1].extend(lj); ent[j][2].extend(-A[k])
                        cnt += 1
        elif typeBc in ("cor", "all", "cor+edge"):
            # Cluster.cpp:846-903 with cornersOnly=true, addDirConstr=false
            corner = np.zeros(3 * mesh.nPoints, dtype=bool)
            corner[mesh.corner_dofs] = True
            for dsub in range(d0, d1):
                ifc = l2g[dsub][glue.mult[l2g[dsub]] > 1]
                for gg in ifc:
                    if not corner[gg]:
                        continue
                    for s in glue.subs_of(gg):
                        if s > dsub and s < d1:
                            ent[dsub][0].append(cnt); ent[dsub][1].append(int(np.searchsorted(l2g[dsub], gg)))
                            ent[dsub][2].append(1.0)
                            ent[s][0].append(cnt); ent[s][1].append(int(np.searchsorted(l2g[s], gg)))
                            ent[s][2].append(-1.0)
                            cnt += 1
                            if full_rank:
                                break
            if typeBc == "cor+edge":
                cnt = _edge_rows(mesh, l2g, glue, d0, d1, ent, cnt, full_rank)
        else:
            raise ValueError(typeBc)
        rows = {d: (np.array(v[0], dtype=np.int64), np.array(v[1], dtype=np.int64), np.array(v[2]))
                for d, v in ent.items()}
        out.append({"nlam": cnt, "rows": rows, "d0": d0, "d1": d1})
    return out


def _edge_rows(mesh, l2g, glue, d0, d1, ent, cnt, full_rank):
    """Averaged edge constraints of typeBc=cor+edge (an extension: the reference has
    only ker/cor/all).  For subdomain d ascending, its 12 box edges in the order axis
    a = x, y, z, then the other two axes (b < c) at (lo, lo), (hi, lo), (lo, hi),
    (hi, hi); the edge's m nodes lie strictly between its end vertices.  For the
    first t > d of the cluster holding all of them (every such t when
    Bc_fullRank=false) one row per component: sum_nodes u_d / m - u_t / m."""
    NX, NY, _ = mesh.NE
    nx1, nxy1 = NX + 1, (NX + 1) * (NY + 1)
    for d in range(d0, d1):
        g = mesh.grid[l2g[d][0::3] // 3]
        lo, hi = g.min(axis=0), g.max(axis=0)
        for a in range(3):
            b, c = [x for x in range(3) if x != a]
            m = int(hi[a] - lo[a]) - 1
            if m < 1:
                continue
            for q in range(4):
                fix = {b: hi[b] if q & 1 else lo[b], c: hi[c] if q & 2 else lo[c]}
                nodes = []
                for e in range(m):
                    p = [0, 0, 0]
                    p[a] = lo[a] + 1 + e
                    p[b], p[c] = fix[b], fix[c]
                    nodes.append(int(p[0] + p[1] * nx1 + p[2] * nxy1))
                owners = set(glue.subs_of(3 * nodes[0]).tolist())
                for nd in nodes[1:]:
                    owners &= set(glue.subs_of(3 * nd).tolist())
                for t in sorted(owners):
                    if t <= d or t >= d1:
                        continue
                    for comp in range(3):
                        dofs = np.array([3 * nd + comp for nd in nodes], dtype=np.int64)
                        ld = np.searchsorted(l2g[d], dofs)
                        lt = np.searchsorted(l2g[t], dofs)
                        ent[d][0].extend([cnt] * m); ent[d][1].extend(ld); ent[d][2].extend([1.0 / m] * m)
                        ent[t][0].extend([cnt] * m); ent[t][1].extend(lt); ent[t][2].extend([-1.0 / m] * m)
                        cnt += 1
                    if full_rank:
                        break
    return cnt
