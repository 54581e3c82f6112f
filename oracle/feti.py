"""HTFETI cluster operators, PCPG and recovery.  TEST INFRASTRUCTURE ONLY.

Restates Cluster::initialization (Cluster.cpp:21-385), the coarse objects
(Cluster.cpp:387-678, 1227-1273, 1332-1352), the operators mult_Kplus_f,
Preconditioning, mult_Ff, mult_Bf/BfT, Projection, scale (Cluster.cpp:961-1360),
pcpg (Cluster.cpp:1363-1458) and the recovery in printVTK (Cluster.cpp:1648-1678).

Choices (identical in the product):
  * the cluster saddle-point matrix: with Ac_clust_explicit=false the two-stage
    Schur-complement solve (Fc^-1, Sc = Gc^T Fc^-1 Gc, (Sc + H H^T)^-1,
    myModul.py:128-135 completing Cluster.cpp:325-342); otherwise, with
    Ac_extended_by_kerGc=true the
    non-singular extended form Ac = [Fc Gc 0; Gc^T 0 H; 0 H^T 0]
    (Cluster.cpp:592-660); with false (the reference default, which factors the
    singular [Fc Gc; Gc^T 0] with its kernel detection, Cluster.cpp:286-320) the
    same matrix regularised along its kernel [0; H], [Fc Gc; Gc^T -sig H H^T], a
    generalised inverse of the singular one.  P F P and u do not depend on which
    consistent generalised inverse is used (pinned: both forms give the same u,
    tests/test_options.py).  A cluster of one subdomain has H = I.
  * Dirichlet preconditioner is the explicit Schur complement (equal to the
    implicit one in exact arithmetic, Matrix.cpp:1853-1870).

Parity: the whole chain is pinned against a global FP64 direct solve
(oracle.direct) in tests/test_oracle_feti.py (displacement <= 1e-8 relative at
eps_iter=1e-10), plus the invariants P^2=P, PG=0, G^T lambda0 = e, symmetry of F
and M, and K K_reg^-1 K = K.
"""
import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from .mesh import Mesh
from .fe import Subdomains
from .decomp import Gluing, build_Bc
from .kernel import kernel_from_K, sset_geometric, projector_pivots, null_pivots, regularize


class _Factor:
    """K_reg^-1 application: dense Cholesky or sparse LU (the reference uses
    PARDISO phase 33, Matrix.cpp:1237-1287)."""

    def __init__(self, Kreg, sparse=False):
        self.sparse = sparse
        if sparse:
            self.lu = spla.splu(sp.csc_matrix(Kreg))
        else:
            self.c = sla.cho_factor(Kreg, lower=True)

    def solve(self, b):
        if self.sparse:
            return self.lu.solve(b)
        return sla.cho_solve(self.c, b)


class HTFETI:
    def __init__(self, o, kernel_mode="geometric", seed=0, sparse=False, subs=None, mesh=None):
        self.o = o
        self.mesh = mesh if mesh is not None else Mesh(o)
        m = self.mesh
        self.subs = subs if subs is not None else Subdomains(m)
        s = self.subs
        self.glue = Gluing(m, s)
        g = self.glue
        self.nsub = m.nsub
        self.nlam = g.nlam
        # ---- kernels and regularised factors (Cluster.cpp:86-109) ----
        self.R, self.defect, self.sset, self.fix, self.Kp = [], [], [], [], []
        for d in range(self.nsub):
            bd = g.Bf[d]["j_col_cmpr"]
            if kernel_mode == "geometric":
                lnodes = s.l2g[d][0::3] // 3
                grid = m.grid[lnodes]
                ss = sset_geometric(grid, bd)
                R, dfc, ss = kernel_from_K(s.K[d], ss)
                fix = ss[projector_pivots(R[ss], dfc)]
            else:
                R, dfc, ss = kernel_from_K(s.K[d], None, seed=seed + d)
                fix = null_pivots(R)
            self.R.append(R)
            self.defect.append(dfc)
            self.sset.append(ss)
            self.fix.append(fix)
            self.Kp.append(_Factor(regularize(s.K[d], fix) if not sparse else
                                   _sparse_reg(s.K[d], fix), sparse))
        # ---- gluing Bc (Cluster.cpp:729-747) ----
        self.Bc = build_Bc(m, s, g, self.R, o["typeBc"], o["Bc_fullRank"] == "true")
        # ---- per-subdomain cluster objects (Cluster.cpp:151-206) ----
        self.BcT, self.KplusBcT, self.Gc, self.bc_rows = [], [], [], []
        for d in range(self.nsub):
            cl = m.cluster_of(d)
            rows, cols, vals = self.Bc[cl]["rows"][d]
            urows = np.unique(rows)
            n = len(s.l2g[d])
            BcT = np.zeros((n, len(urows)))
            np.add.at(BcT, (cols, np.searchsorted(urows, rows)), vals)
            self.BcT.append(BcT)
            self.bc_rows.append(urows)
            self.KplusBcT.append(self.Kp[d].solve(BcT) if len(urows) else np.zeros((n, 0)))
            self.Gc.append(-(BcT.T @ self.R[d]))
        # ---- cluster coarse problems (Cluster.cpp:209-324) ----
        self.clusters = []
        for cl in range(m.ncl):
            d0, d1 = self.Bc[cl]["d0"], self.Bc[cl]["d1"]
            nlc = self.Bc[cl]["nlam"]
            nr = sum(self.R[d].shape[1] for d in range(d0, d1))
            Fc = np.zeros((nlc, nlc))
            Gc = np.zeros((nlc, nr))
            off = 0
            roff = []
            for d in range(d0, d1):
                ur = self.bc_rows[d]
                Fc[np.ix_(ur, ur)] += self.BcT[d].T @ self.KplusBcT[d]
                Gc[ur, off:off + self.R[d].shape[1]] += self.Gc[d]
                roff.append(off)
                off += self.R[d].shape[1]
            GtG = Gc.T @ Gc
            if d1 - d0 == 1:
                # one subdomain: no cluster constraints, Gc^T Gc = 0 and its kernel is
                # the whole 6-dimensional alpha space
                H = np.eye(nr)
            else:
                H, dH, _ = kernel_from_K(GtG, None, seed=seed + 100000 + cl)
            nh = H.shape[1]
            two_stage = None
            if o.get("Ac_clust_explicit", "true") == "false":
                # the Sc branch (Cluster.cpp:325-342: Fc factorised, Sc = Gc^T Fc^-1 Gc)
                # completed as the two-stage coarse solve of the Python HTFETI
                # (myModul.py:128-135, 500-510): iF0 = Fc^-1, iS0 = (Sc + H H^T)^-1,
                # beta = iS0 (Gc^T iF0 d0 - e0), mu = iF0 (d0 - Gc beta)
                iF0 = sla.cho_factor(Fc, lower=True)
                Sc = Gc.T @ sla.cho_solve(iF0, Gc)
                two_stage = (iF0, sla.cho_factor(Sc + H @ H.T, lower=True), Sc)
            if o.get("Ac_extended_by_kerGc", "true") == "true":
                # extended non-singular form (Cluster.cpp:592-660)
                Ac = np.zeros((nlc + nr + nh, nlc + nr + nh))
                Ac[:nlc, :nlc] = Fc
                Ac[:nlc, nlc:nlc + nr] = Gc
                Ac[nlc:nlc + nr, :nlc] = Gc.T
                Ac[nlc:nlc + nr, nlc + nr:] = H
                Ac[nlc + nr:, nlc:nlc + nr] = H.T
            else:
                # the singular Ac = [Fc Gc; Gc^T 0] with its kernel [0; H] regularised:
                # A_r = Ac - sig [0 0; 0 H H^T] is a generalised inverse of Ac (A_r [0; H y]
                # = [0; -sig H y]), sig = (tr(Gc^T Gc)/nr) / (tr(Fc)/nlc)
                tf, tg = np.trace(Fc), np.trace(GtG)
                sig = (tg / nr) / (tf / nlc) if (nlc > 0 and tf > 0) else 1.0
                Ac = np.zeros((nlc + nr + nh, nlc + nr + nh))
                Ac[:nlc, :nlc] = Fc
                Ac[:nlc, nlc:nlc + nr] = Gc
                Ac[nlc:nlc + nr, :nlc] = Gc.T
                Ac[nlc:nlc + nr, nlc:nlc + nr] = -sig * (H @ H.T)
                Ac[nlc + nr:, nlc + nr:] = np.eye(nh)  # decoupled rows: multiplier 0
            self.clusters.append({"d0": d0, "d1": d1, "nlc": nlc, "nr": nr, "roff": roff, "H": H,
                                  "Fc": Fc, "Gc": Gc, "Ac_lu": sla.lu_factor(Ac), "nh": nh,
                                  "Ac": Ac, "two_stage": two_stage})
        # ---- FETI coarse: Rf, Gf, GfTGf (Cluster.cpp:1227-1273, 1332-1352) ----
        ncol = sum(c["nh"] for c in self.clusters)
        self.ncol_G = ncol
        self.Rf = []
        coff = np.cumsum([0] + [c["nh"] for c in self.clusters])
        self.coff = coff
        Gf = np.zeros((self.nlam, ncol))
        for d in range(self.nsub):
            cl = m.cluster_of(d)
            c = self.clusters[cl]
            ro = c["roff"][d - c["d0"]]
            Hd = c["H"][ro:ro + self.R[d].shape[1], :]
            Rf = np.zeros((len(s.l2g[d]), ncol))
            Rf[:, coff[cl]:coff[cl + 1]] = self.R[d] @ Hd
            self.Rf.append(Rf)
            bf = g.Bf[d]
            np.add.at(Gf, bf["rows"], bf["val"][:, None] * Rf[bf["col"], :])
        self.Gf = Gf
        self.GfTGf = Gf.T @ Gf
        c = sla.cho_factor(self.GfTGf, lower=True)
        inv = sla.cho_solve(c, np.eye(ncol))
        self.invGfTGf = 0.5 * (inv + inv.T)
        # ---- preconditioner (Cluster.cpp:140-148, Matrix.cpp:1708-1871) ----
        # Dirichlet_explicit and Dirichlet_implicit build the same operator S_bb
        # (explicit vs implicit application only); every other value applies K
        # (the lumped branch, Cluster.cpp:1104-1107)
        self.prec = o["preconditioner"]
        dirichlet = self.prec in ("Dirichlet_explicit", "Dirichlet_implicit")
        self.S = []
        for d in range(self.nsub):
            b = g.Bf[d]["j_col_cmpr"]
            K = s.K[d]
            if dirichlet:
                r = np.setdiff1d(np.arange(K.shape[0]), b)
                Kd = K.toarray() if hasattr(K, "toarray") else K
                Kbb = Kd[np.ix_(b, b)]
                Krb = Kd[np.ix_(r, b)]
                S = Kbb - Krb.T @ sla.cho_solve(sla.cho_factor(Kd[np.ix_(r, r)], lower=True), Krb)
                self.S.append(0.5 * (S + S.T))
            else:
                self.S.append(None)

    # ------------------------------------------------------------------ Bf
    def mult_BfT(self, lam):
        """x_d = Bf_d^T lam (Cluster.cpp:1210-1224)."""
        out = []
        for d in range(self.nsub):
            bf = self.glue.Bf[d]
            x = np.zeros(len(self.subs.l2g[d]))
            np.add.at(x, bf["col"], bf["val"] * lam[bf["rows"]])
            out.append(x)
        return out

    def mult_Bf(self, xs):
        """lam = sum_d Bf_d x_d (Cluster.cpp:1193-1208)."""
        lam = np.zeros(self.nlam)
        for d in range(self.nsub):
            bf = self.glue.Bf[d]
            np.add.at(lam, bf["rows"], bf["val"] * xs[d][bf["col"]])
        return lam

    # ------------------------------------------------------------ K_c^+
    def _cluster_correction(self, xs, ys=None):
        """Shared part of mult_Kplus_f / mult_Ff (Cluster.cpp:975-1040, 1129-1181):
        returns per subdomain R_d alpha_d - KplusBcT_d lamc_d."""
        out = [None] * self.nsub
        for cl, c in enumerate(self.clusters):
            nlc, nr, nh = c["nlc"], c["nr"], c["nh"]
            gc = np.zeros(nlc + nr + nh)
            for d in range(c["d0"], c["d1"]):
                gc[self.bc_rows[d]] += self.KplusBcT[d].T @ xs[d]
                ro = c["roff"][d - c["d0"]]
                gc[nlc + ro:nlc + ro + self.R[d].shape[1]] = -(self.R[d].T @ xs[d])
            if c["two_stage"] is not None:
                iF0, iS0, _ = c["two_stage"]
                d0v, e0v = gc[:nlc], gc[nlc:nlc + nr]
                beta = sla.cho_solve(iS0, c["Gc"].T @ sla.cho_solve(iF0, d0v) - e0v)
                mu = sla.cho_solve(iF0, d0v - c["Gc"] @ beta)
                la = np.concatenate([mu, beta, np.zeros(nh)])
            else:
                la = sla.lu_solve(c["Ac_lu"], gc)
            for d in range(c["d0"], c["d1"]):
                ro = c["roff"][d - c["d0"]]
                alpha = la[nlc + ro:nlc + ro + self.R[d].shape[1]]
                out[d] = self.R[d] @ alpha - self.KplusBcT[d] @ la[self.bc_rows[d]]
        return out

    def mult_Kplus_f(self, fs):
        """x = K_c^+ f (Cluster.cpp:967-1040)."""
        corr = self._cluster_correction(fs)
        return [self.Kp[d].solve(fs[d]) + corr[d] for d in range(self.nsub)]

    def mult_Ff(self, w):
        """F w = Bf K_c^+ Bf^T w (Cluster.cpp:1114-1188)."""
        xs = self.mult_BfT(w)
        return self.mult_Bf(self.mult_Kplus_f(xs))

    # ------------------------------------------------------------ projector
    def projection(self, x):
        """Px = x - Gf inv(GfTGf) Gf^T x; alpha = -inv(GfTGf) Gf^T x
        (Cluster.cpp:1314-1329)."""
        a = self.invGfTGf @ (self.Gf.T @ x)
        return x - self.Gf @ a, -a

    def scale(self, x):
        return x * self.glue.weight  # Cluster.cpp:1356-1360

    def preconditioning(self, w):
        """M w = W Bf S Bf^T W w (Cluster.cpp:1043-1112)."""
        xs = self.mult_BfT(self.scale(w))
        ys = []
        for d in range(self.nsub):
            y = np.zeros_like(xs[d])
            if self.S[d] is not None:
                b = self.glue.Bf[d]["j_col_cmpr"]
                y[b] = self.S[d] @ xs[d][b]
            else:
                y = self.subs.K[d] @ xs[d]
            ys.append(y)
        return self.scale(self.mult_Bf(ys))

    # ------------------------------------------------------------ PCPG
    def rhs(self):
        return self.subs.f

    def pcpg(self, f=None, eps=None, max_iter=None, history=False):
        """Cluster::pcpg (Cluster.cpp:1363-1458).  Returns dict with lambda,
        alpha, iterations (number of F applications inside the loop), u."""
        o = self.o
        eps = float(o["eps_iter"]) if eps is None else eps
        max_iter = int(o["max_iter"]) if max_iter is None else max_iter
        f = self.rhs() if f is None else f
        xx = self.mult_Kplus_f(f)
        d_rhs = self.mult_Bf(xx)
        e = np.zeros(self.ncol_G)
        for d in range(self.nsub):
            e += self.Rf[d].T @ f[d]
        lam = self.Gf @ (self.invGfTGf @ e)
        g = self.mult_Ff(lam) - d_rhs
        Pg, alpha = self.projection(g)
        z = self.preconditioning(Pg)
        Pz, _ = self.projection(z)
        gPz = float(np.dot(g, Pz))
        norm0 = np.sqrt(gPz)
        w = Pz.copy()
        hist = [1.0]
        it = 0
        for it in range(max_iter + 1):
            if it == max_iter or np.sqrt(gPz) < eps * norm0:
                break
            Fw = self.mult_Ff(w)
            wFw = float(np.dot(w, Fw))
            rho = -gPz / wFw
            lam = lam + rho * w
            g = g + rho * Fw
            Pg, alpha = self.projection(g)
            z = self.preconditioning(Pg)
            Pz, _ = self.projection(z)
            gPz_prev = gPz
            gPz = float(np.dot(g, Pz))
            gamma = gPz / gPz_prev
            w = Pz + gamma * w
            hist.append(np.sqrt(gPz) / norm0)
        u = self.recover(f, lam, alpha)
        out = {"lambda": lam, "alpha": alpha, "iterations": it, "u": u,
               "residual": np.sqrt(max(gPz, 0.0)) / norm0, "e": e, "d": d_rhs}
        if history:
            out["history"] = hist
        return out

    def recover(self, f, lam, alpha):
        """u = K_c^+(f - Bf^T lam) - Rf alpha, scattered by l2g, later subdomains
        overwrite shared DOFs (Cluster.cpp:1648-1678)."""
        bt = self.mult_BfT(lam)
        yy = [f[d] - bt[d] for d in range(self.nsub)]
        xx = self.mult_Kplus_f(yy)
        u = np.zeros(3 * self.mesh.nPoints)
        for d in range(self.nsub):
            u[self.subs.l2g[d]] = xx[d] - self.Rf[d] @ alpha
        return u


def _sparse_reg(K, fix):
    K = sp.csr_matrix(K, copy=True)
    dg = K.diagonal()
    K = K + sp.csr_matrix((dg[fix], (fix, fix)), shape=K.shape)
    return K
