"""CPU baseline: the C++ FP64 HTFETI restatement (oracle/cpu/htfeti_cpu.cpp) driven
from the oracle's index structures.  TEST INFRASTRUCTURE ONLY: bench.py times it
as the CPU baseline and tests/test_cpu_port.py checks it against the Python oracle
and the golden fixtures; nothing in hfls_amd/ imports it.

The C++ side follows the reference's PARDISO path (htfeti.cpp:27-62: analytic
kernel, preconditioner=Dirichlet_implicit): hex8 assembly, analytic R
(Data.cpp:61-109), fixing DOFs (Matrix.cpp:835-884), K_reg factorisation, K_rr
factorisation, cluster objects and the extended cluster matrix
(Cluster.cpp:151-206, 592-660), Gf and (Gf^T Gf)^-1, PCPG with an implicit K_c^+
per iteration (Cluster.cpp:1363-1458) and the recovery (Cluster.cpp:1648-1678),
with OpenMP over subdomains.  The integer structures (numbering, l2g, Bf rows and
weights, Bc) come from oracle.mesh / oracle.decomp (Mesh.cpp:69-289,
Cluster.cpp:682-959) and are not part of the timed set-up.
"""
import ctypes
import os
import subprocess

import numpy as np

from .mesh import Mesh
from .decomp import Gluing, build_Bc

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpu", "htfeti_cpu.cpp")
LIB = os.path.join(HERE, "cpu", "libhtfeti_cpu.so")


def build():
    """g++ -O3 (AVX2 + FMA, OpenMP) of the C++ restatement into oracle/cpu/."""
    tmp = LIB + ".tmp"
    cmd = ["g++", "-O3", "-mavx2", "-mfma", "-fopenmp", "-std=c++17", "-shared", "-fPIC", SRC, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("CPU baseline build failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    return LIB


class _Desc(ctypes.Structure):
    _fields_ = [("nsub", ctypes.c_int), ("ncl", ctypes.c_int), ("nlam", ctypes.c_int64), ("ndof", ctypes.c_int64),
                ("nu", ctypes.c_double), ("acc", ctypes.c_double * 3)] + \
               [(n, ctypes.c_void_p) for n in ("n_off", "l2g", "coords", "e_off", "conn", "Eel", "bf_off", "bf_row",
                                                "bf_col", "bf_val", "weight", "bc_off", "bc_row", "bc_col", "bc_val",
                                                "cl_d0", "cl_d1", "cl_nlc")] + [("lumped", ctypes.c_int)]


class _Times(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("assemble", "kernel_factor", "dirichlet", "cluster", "coarse", "total")]


class _STimes(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("init", "pcpg", "recover", "total")]


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("CPU baseline library missing: run __graft_entry__.build() (or oracle.cpu_port.build())")
        _lib = ctypes.CDLL(LIB)
        _lib.htc_setup.restype = ctypes.c_void_p
        _lib.htc_setup.argtypes = [ctypes.POINTER(_Desc), ctypes.c_int, ctypes.POINTER(_Times)]
        _lib.htc_solve.restype = ctypes.c_int
        _lib.htc_solve.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(_STimes)]
        _lib.htc_last_error.restype = ctypes.c_char_p
        _lib.htc_free.argtypes = [ctypes.c_void_p]
        _lib.htc_body_load.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _lib.htc_fix.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    return _lib


class _Lite:
    """l2g and element lists per subdomain (Data.cpp:111-144), without K."""

    def __init__(self, mesh):
        order = np.argsort(mesh.part, kind="stable")  # ascending element index within each subdomain
        counts = np.bincount(mesh.part, minlength=mesh.nsub)
        starts = np.concatenate([[0], np.cumsum(counts)])
        self.elems, self.l2g, self.conn = [], [], []
        for d in range(mesh.nsub):
            el = order[starts[d]:starts[d + 1]]
            nodes = mesh.elements[el]
            un = np.unique(nodes)
            self.elems.append(el)
            self.l2g.append((3 * un[:, None] + np.arange(3)[None, :]).ravel())
            self.conn.append(np.searchsorted(un, nodes).astype(np.int32))


def analytic_R(mesh, l2g):
    """Data::create_analytic_ker_K (Data.cpp:61-109): rigid-body modes from the
    global node coordinates, each column normalised."""
    nodes = l2g[0::3] // 3
    x, y, z = mesh.points[nodes].T
    n = len(l2g)
    R = np.zeros((n, 6))
    R[0::3, 0] = 1
    R[1::3, 1] = 1
    R[2::3, 2] = 1
    R[0::3, 3], R[1::3, 3] = -y, x
    R[0::3, 4], R[2::3, 4] = -z, x
    R[1::3, 5], R[2::3, 5] = -z, y
    return R / np.linalg.norm(R, axis=0)


class CpuHTFETI:
    """Set-up on `threads` OpenMP threads (wall times per phase in `setup_times`);
    solve(f) runs PCPG and the recovery for one load case."""

    def __init__(self, o, threads=1):
        self.o = o
        m = self.mesh = Mesh(o)
        lite = self.lite = _Lite(m)
        g = self.glue = Gluing(m, lite)
        R = [analytic_R(m, l2g) for l2g in lite.l2g]
        Bc = build_Bc(m, lite, g, R, o["typeBc"], o.get("Bc_fullRank", "true") == "true")
        N = m.nsub
        self.nsub, self.nlam = N, g.nlam
        self.npacked = sum(len(x) for x in lite.l2g)
        a = self._arrays = {}
        a["n_off"] = np.concatenate([[0], np.cumsum([len(x) for x in lite.l2g])]).astype(np.int64)
        a["l2g"] = np.concatenate(lite.l2g).astype(np.int64)
        a["coords"] = np.ascontiguousarray(np.concatenate([m.points[x[0::3] // 3].ravel() for x in lite.l2g]))
        a["e_off"] = np.concatenate([[0], np.cumsum([len(e) for e in lite.elems])]).astype(np.int64)
        a["conn"] = np.ascontiguousarray(np.concatenate(lite.conn).astype(np.int32))
        a["Eel"] = np.ascontiguousarray(np.concatenate([m.young_per_element()[e] for e in lite.elems]), dtype=np.float64)
        a["bf_off"] = np.concatenate([[0], np.cumsum([len(g.Bf[d]["rows"]) for d in range(N)])]).astype(np.int64)
        a["bf_row"] = np.concatenate([g.Bf[d]["rows"] for d in range(N)]).astype(np.int64)
        a["bf_col"] = np.concatenate([g.Bf[d]["col"] for d in range(N)]).astype(np.int32)
        a["bf_val"] = np.concatenate([g.Bf[d]["val"] for d in range(N)]).astype(np.float64)
        a["weight"] = np.ascontiguousarray(g.weight, dtype=np.float64)
        rows, cols, vals = [], [], []
        for cl in Bc:
            for d in range(cl["d0"], cl["d1"]):
                r, c, v = cl["rows"][d]
                rows.append(r), cols.append(c), vals.append(v)
        a["bc_off"] = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int64)
        a["bc_row"] = np.concatenate(rows).astype(np.int32)
        a["bc_col"] = np.concatenate(cols).astype(np.int32)
        a["bc_val"] = np.concatenate(vals).astype(np.float64)
        a["cl_d0"] = np.array([c["d0"] for c in Bc], dtype=np.int32)
        a["cl_d1"] = np.array([c["d1"] for c in Bc], dtype=np.int32)
        a["cl_nlc"] = np.array([c["nlam"] for c in Bc], dtype=np.int32)
        d = _Desc()
        d.nsub, d.ncl, d.nlam, d.ndof = N, m.ncl, g.nlam, 3 * m.nPoints
        d.nu = float(o["poissons_ratio"])
        acc = {"reference": (0.0, 0.0, 2.0 * 98.10)}.get(o["load"], (0.0, 0.0, float(o.get("body_force", "-1.0"))))
        d.acc = (ctypes.c_double * 3)(*acc)
        for k, v in a.items():
            setattr(d, k, v.ctypes.data)
        prec = o.get("preconditioner", "Dirichlet_implicit")
        d.lumped = 0 if prec in ("Dirichlet_implicit", "Dirichlet_explicit") else 1
        self._desc = d
        t = _Times()
        self._h = _load().htc_setup(ctypes.byref(d), int(threads), ctypes.byref(t))
        if not self._h:
            raise RuntimeError("CPU baseline set-up: " + _load().htc_last_error().decode())
        self.setup_times = {k: getattr(t, k) for k, _ in _Times._fields_}

    def close(self):
        if getattr(self, "_h", None):
            _load().htc_free(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def body_load(self):
        f = np.zeros(self.npacked)
        _load().htc_body_load(self._h, f.ctypes.data)
        return f

    def fixing_dofs(self, d):
        out = np.zeros(6, dtype=np.int32)
        _load().htc_fix(self._h, int(d), out.ctypes.data)
        return out

    def solve(self, f=None, eps=1e-8, max_iter=500, threads=1):
        """One load case (packed per subdomain in l2g order, None: the body load).
        Returns u (global), iterations, residual and the wall times."""
        if f is not None:
            f = np.ascontiguousarray(f, dtype=np.float64)
            if f.shape != (self.npacked,):
                raise ValueError("f must have %d entries" % self.npacked)
        u = np.zeros(3 * self.mesh.nPoints)
        it, res, t = ctypes.c_int(0), ctypes.c_double(0), _STimes()
        rc = _load().htc_solve(self._h, None if f is None else f.ctypes.data, float(eps), int(max_iter), int(threads),
                               u.ctypes.data, ctypes.byref(it), ctypes.byref(res), ctypes.byref(t))
        if rc != 0:
            raise RuntimeError("CPU baseline solve: " + _load().htc_last_error().decode())
        return {"u": u, "iterations": it.value, "residual": res.value,
                "times": {k: getattr(t, k) for k, _ in _STimes._fields_}}
