"""MI355X-native HTFETI hot path: Python host mirror of the reference's
Driver/Cluster/Solver entry points over the C ABI of libhfls.so (include/hfls_gpu.h).

The native library is mandatory: every call goes through the in-tree
``libhfls.so`` (gfx950 HIP kernels); if it is missing the import of the binding
raises instead of falling back to any CPU path.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhfls.so")

EXPORTS = [
    "hfls_last_error", "hfls_problem_create", "hfls_problem_destroy", "hfls_problem_sizes",
    "hfls_problem_sub_info", "hfls_problem_sub_bf", "hfls_problem_sub_bdofs", "hfls_problem_sub_sset",
    "hfls_problem_sub_l2g", "hfls_problem_sub_young", "hfls_problem_rows", "hfls_problem_dirichlet",
    "hfls_problem_corners", "hfls_mesh_create", "hfls_mesh_destroy", "hfls_mesh_sizes",
    "hfls_mesh_setup_times", "hfls_mesh_kernel", "hfls_mesh_load", "hfls_apply_F", "hfls_project",
    "hfls_precondition", "hfls_kplus", "hfls_solve", "hfls_time_stream", "hfls_solve_block",
    "hfls_apply_block", "hfls_problem_traction", "hfls_mesh_traction", "hfls_server_create",
    "hfls_server_destroy", "hfls_server_sizes", "hfls_server_submit", "hfls_server_step", "hfls_server_state",
    "hfls_server_result", "hfls_server_stats", "hfls_mesh_configure",
    "hfls_mesh_tuning", "hfls_problem_write_vtk", "hfls_mesh_write_vtk", "hfls_solve_blocks",
    "hfls_server_cancel", "hfls_mult_Gf", "hfls_mult_GfT", "hfls_mult_RfT", "hfls_recover", "hfls_mesh_rf",
    "hfls_mesh_counts", "hfls_project_noalloc", "hfls_mesh_save", "hfls_mesh_restore", "hfls_server_host_cache",
    "hfls_server_cache_stats", "hfls_host_alloc", "hfls_host_free",
]

DEFAULTS = {
    "Nx": "2", "Ny": "2", "Nz": "2", "nx": "5", "ny": "5", "nz": "5",
    "Cx": "1", "Cy": "1", "Cz": "1",
    "young_modulus": "1000", "ratio_mat": "1e4", "poissons_ratio": "0.3",
    "material": "reference", "mat_sigma": "1.5", "mat_contrast": "1e3", "seed": "0",
    "bc": "symmetry", "load": "reference", "body_force": "-1.0", "density": "2.0",
    "typeBc": "ker", "Bc_fullRank": "true", "preconditioner": "Dirichlet_explicit",
    "eps_iter": "1e-6", "max_iter": "200", "Ac_extended_by_kerGc": "true",
    "slabs_x": "", "slabs_y": "", "slabs_z": "",
    "linear_solver": "hip", "vtkWithinIter": "false", "print_matrices": "0",
    "precision": "fp64", "nrhs": "1", "device": "0", "refine_tol": "1e-4", "slab_choices": "", "block_gemm": "f16x2",
}


class HflsError(RuntimeError):
    pass


class Stats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("converged", ctypes.c_int), ("residual", ctypes.c_double),
                ("solve_ms", ctypes.c_double), ("pcpg_ms", ctypes.c_double), ("pg_rel", ctypes.c_double)]


class CaseStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("status", ctypes.c_int), ("residual", ctypes.c_double)]


class BlockStats(ctypes.Structure):
    _fields_ = [("solve_ms", ctypes.c_double), ("init_ms", ctypes.c_double), ("pcpg_ms", ctypes.c_double),
                ("recover_ms", ctypes.c_double), ("sweeps", ctypes.c_int), ("cert_steps", ctypes.c_int)]


_lib = None


def load():
    """Load libhfls.so (built by hfls_amd.build); raise if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HflsError("libhfls.so not built: run python -m hfls_amd.build")
    lib = ctypes.CDLL(LIB_PATH)
    lib.hfls_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def _opts(o):
    o = {k: ("" if v is None else (",".join(str(int(x)) for x in v) if isinstance(v, (list, tuple)) else str(v)))
         for k, v in o.items()}
    keys = (ctypes.c_char_p * len(o))(*[k.encode() for k in o])
    vals = (ctypes.c_char_p * len(o))(*[v.encode() for v in o.values()])
    return len(o), keys, vals


def _check(rc):
    if rc != 0:
        raise HflsError(load().hfls_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class _HostAlloc:
    """Owner of one hfls_host_alloc block; frees it when the last array over it goes."""

    def __init__(self, addr):
        self.addr = addr

    def __del__(self):
        if self.addr and _lib is not None:
            _lib.hfls_host_free(ctypes.c_void_p(self.addr))
            self.addr = None


def pinned_empty(shape, dtype=np.float64):
    """Uninitialised array in page-locked host memory (hfls_host_alloc): loads and
    results in such arrays move to and from the device as DMA at the link rate.
    For load and result buffers that are reused across solves (pcpg_block's F and
    out=)."""
    shape = tuple(int(x) for x in np.atleast_1d(shape))
    n = int(np.prod(shape)) * np.dtype(dtype).itemsize
    if n <= 0:
        return np.empty(shape, dtype=dtype)
    p = ctypes.c_void_p()
    _check(load().hfls_host_alloc(ctypes.c_int64(n), ctypes.byref(p)))
    buf = (ctypes.c_char * n).from_address(p.value)
    buf._owner = _HostAlloc(p.value)  # every view's base chain ends at buf
    return np.frombuffer(buf, dtype=dtype).reshape(shape)


def _vec(x, n, name):
    """x as a contiguous float64 vector of exactly n entries (the C ABI reads n)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    if x.ndim != 1 or x.shape[0] != n:
        raise ValueError("%s must have shape (%d,), got %s" % (name, n, x.shape))
    return x


def _rows(X, n, name):
    """X as a contiguous float64 (k, n) block with k >= 1 rows."""
    X = np.ascontiguousarray(np.atleast_2d(X), dtype=np.float64)
    if X.ndim != 2 or X.shape[1] != n or X.shape[0] < 1:
        raise ValueError("%s must have shape (k >= 1, %d), got %s" % (name, n, X.shape))
    return X


class Problem:
    """Host-side mesh, numbering and gluing (Mesh.cpp:69-289, Data.cpp:111-144,
    Cluster.cpp:682-959) computed by the native library."""

    def __init__(self, **opts):
        lib = load()
        n, k, v = _opts(opts)
        self._h = ctypes.c_void_p()
        _check(lib.hfls_problem_create(n, k, v, ctypes.byref(self._h)))
        s = np.zeros(8, dtype=np.int64)
        _check(lib.hfls_problem_sizes(self._h, _ptr(s)))
        (self.nsub, self.nlam, self.n_glue, self.nPoints, self.ncl, self.nsub_cl, self.ndir,
         self.ncorner) = (int(x) for x in s)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.hfls_problem_destroy(self._h)
            self._h = None

    def sub_info(self, d):
        info = np.zeros(8, dtype=np.int32)
        _check(load().hfls_problem_sub_info(self._h, d, _ptr(info)))
        return dict(zip(["n", "n_r", "n_s", "n_bs", "n_ss", "cluster", "nbf"], (int(x) for x in info[:7])))

    def write_vtk(self, u, folder, iter=-1):
        """Legacy ASCII VTK of the displacements u (3 * nPoints) as <folder>/box.vtk
        (iter < 0) or <folder>/box<iter>.vtk (Mesh::SaveVTK, Mesh.cpp:292-351)."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.size != 3 * self.nPoints:
            raise ValueError("u must hold 3 * nPoints values")
        _check(load().hfls_problem_write_vtk(self._h, _ptr(u), str(folder).encode(), int(iter)))

    def sub_bf(self, d):
        m = self.sub_info(d)["nbf"]
        rows, dofs, signs = (np.zeros(m, dtype=np.int32) for _ in range(3))
        _check(load().hfls_problem_sub_bf(self._h, d, _ptr(rows), _ptr(dofs), _ptr(signs)))
        return rows, dofs, signs

    def sub_bdofs(self, d):
        i = self.sub_info(d)
        out = np.zeros(i["n_s"], dtype=np.int32)
        _check(load().hfls_problem_sub_bdofs(self._h, d, _ptr(out)))
        return out

    def sub_sset(self, d):
        i = self.sub_info(d)
        out = np.zeros(i["n_ss"], dtype=np.int32)
        _check(load().hfls_problem_sub_sset(self._h, d, _ptr(out)))
        return out

    def sub_l2g(self, d):
        out = np.zeros(self.sub_info(d)["n"], dtype=np.int64)
        _check(load().hfls_problem_sub_l2g(self._h, d, _ptr(out)))
        return out

    def sub_young(self, d):
        e = ctypes.c_double()
        _check(load().hfls_problem_sub_young(self._h, d, ctypes.byref(e)))
        return e.value

    def rows(self):
        a, b = np.zeros(self.nlam, dtype=np.int32), np.zeros(self.nlam, dtype=np.int32)
        dof, w = np.zeros(self.nlam, dtype=np.int64), np.zeros(self.nlam)
        _check(load().hfls_problem_rows(self._h, _ptr(a), _ptr(b), _ptr(dof), _ptr(w)))
        return a, b, dof, w

    def dirichlet(self):
        out = np.zeros(self.ndir, dtype=np.int64)
        _check(load().hfls_problem_dirichlet(self._h, _ptr(out)))
        return out

    def corners(self):
        out = np.zeros(self.ncorner, dtype=np.int64)
        _check(load().hfls_problem_corners(self._h, _ptr(out)))
        return out

    def npacked(self):
        return sum(self.sub_info(d)["n"] for d in range(self.nsub))

    def traction(self, seed, case0, k):
        """k seeded smoothed tractions on the x = 1 face, packed local loads (k, npacked)."""
        out = np.zeros((k, self.npacked()))
        _check(load().hfls_problem_traction(self._h, ctypes.c_uint64(seed), int(case0), int(k), _ptr(out)))
        return out


class Mesh:
    """Device-resident, set-up HTFETI mesh (one GPU).  Mirrors the operator
    methods of the reference Cluster (Cluster.hpp:106-121)."""

    def __init__(self, device=0, **opts):
        lib = load()
        n, k, v = _opts(opts)
        self._h = ctypes.c_void_p()
        _check(lib.hfls_mesh_create(int(device), n, k, v, ctypes.byref(self._h)))
        self.restore_ms = None
        self._read_sizes()

    @classmethod
    def restore(cls, path, device=0):
        """A mesh from the set-up state saved by save() (in this or another process),
        without the factorisation; bit-identical results to the saved mesh.
        ``restore_ms`` is the load time, ``setup_ms`` the original set-up's."""
        lib = load()
        self = cls.__new__(cls)
        self._h = ctypes.c_void_p()
        ms = ctypes.c_double()
        _check(lib.hfls_mesh_restore(int(device), os.fsencode(path), ctypes.byref(self._h), ctypes.byref(ms)))
        self.restore_ms = ms.value
        self._read_sizes()
        return self

    def save(self, path):
        """Write the set-up state (options, factors, K+_bb and S_i tiles, kernels,
        gluing, coarse inverses) to ``path`` (through path + ".part")."""
        _check(load().hfls_mesh_save(self._h, os.fsencode(path)))

    def _read_sizes(self):
        lib = load()
        s = np.zeros(6, dtype=np.int64)
        _check(lib.hfls_mesh_sizes(self._h, _ptr(s)))
        self.nsub, self.nlam, self.ncol, self.npacked, self.nPoints, self.device_bytes = (int(x) for x in s)
        t = np.zeros(5)
        _check(lib.hfls_mesh_setup_times(self._h, _ptr(t)))
        self.setup_ms, self.factor_ms = float(t[0]), float(t[1])
        self.setup_phases_ms = {"index": float(t[2]), "assemble": float(t[3]), "factor": float(t[1]),
                                "coarse": float(t[4])}

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.hfls_mesh_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def configure(self, **settings):
        """Runtime settings: hip_graph=True/False, refine_tol=..., poll_chunk=...,
        gemm_rows=1|2|"auto", gemm_kernel="bands"|"rows"."""
        for k, v in settings.items():
            if isinstance(v, bool):
                v = "1" if v else "0"
            _check(load().hfls_mesh_configure(self._h, k.encode(), str(v).encode()))

    def tuning(self):
        """Autotuner result of the block GEMM: chosen row tiles per workgroup and the
        measured ms of the candidates (after the first block call); ms[3] is the band
        form of the split-FP16 GEMM (setting gemm_kernel)."""
        out = np.zeros(4)
        _check(load().hfls_mesh_tuning(self._h, _ptr(out)))
        return {"row_tiles": int(out[0]), "ms": {1: float(out[1]), 2: float(out[2]), 3: float(out[3])}}

    def write_vtk(self, u, folder, iter=-1):
        """Legacy ASCII VTK of the displacements u (3 * nPoints), as Problem.write_vtk."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.size != 3 * self.nPoints:
            raise ValueError("u must hold 3 * nPoints values")
        _check(load().hfls_mesh_write_vtk(self._h, _ptr(u), str(folder).encode(), int(iter)))

    def kernel(self):
        R = np.zeros(6 * self.npacked)
        fix = np.zeros(6 * self.nsub, dtype=np.int32)
        dfc = np.zeros(self.nsub, dtype=np.int32)
        _check(load().hfls_mesh_kernel(self._h, _ptr(R), _ptr(fix), _ptr(dfc)))
        return R, fix, dfc

    def load_vector(self):
        f = np.zeros(self.npacked)
        _check(load().hfls_mesh_load(self._h, _ptr(f)))
        return f

    def mult_Ff(self, w):
        w = _vec(w, self.nlam, "w")
        out = np.zeros(self.nlam)
        _check(load().hfls_apply_F(self._h, _ptr(w), _ptr(out)))
        return out

    def projection(self, x):
        x = _vec(x, self.nlam, "x")
        Px, alpha = np.zeros(self.nlam), np.zeros(self.ncol)
        _check(load().hfls_project(self._h, _ptr(x), _ptr(Px), _ptr(alpha)))
        return Px, alpha

    def preconditioning(self, x):
        x = _vec(x, self.nlam, "x")
        out = np.zeros(self.nlam)
        _check(load().hfls_precondition(self._h, _ptr(x), _ptr(out)))
        return out

    def mult_Kplus_f(self, f_packed):
        f = _vec(f_packed, self.npacked, "f_packed")
        out = np.zeros(self.npacked)
        _check(load().hfls_kplus(self._h, _ptr(f), _ptr(out)))
        return out

    def mult_Gf(self, alpha):
        """Gf alpha (Cluster.cpp:1275-1283); alpha has 6 entries per cluster."""
        alpha = _vec(alpha, self.ncol, "alpha")
        out = np.zeros(self.nlam)
        _check(load().hfls_mult_Gf(self._h, _ptr(alpha), _ptr(out)))
        return out

    def mult_GfT(self, lam):
        """Gf^T lambda (Cluster.cpp:1285-1293)."""
        lam = _vec(lam, self.nlam, "lambda")
        out = np.zeros(self.ncol)
        _check(load().hfls_mult_GfT(self._h, _ptr(lam), _ptr(out)))
        return out

    def mult_RfT(self, f_packed):
        """sum_d Rf_d^T f_d (Cluster.cpp:1295-1312), f packed local."""
        f = _vec(f_packed, self.npacked, "f_packed")
        out = np.zeros(self.ncol)
        _check(load().hfls_mult_RfT(self._h, _ptr(f), _ptr(out)))
        return out

    def recover(self, f_packed, lam, alpha):
        """Recovery of printVTK (Cluster.cpp:1648-1678): (x packed local, u global)
        with x_d = K_c^+ (f - Bf^T lambda)_d - Rf_d alpha, alpha as from projection()."""
        f = _vec(f_packed, self.npacked, "f_packed")
        lam = _vec(lam, self.nlam, "lambda")
        alpha = _vec(alpha, self.ncol, "alpha")
        x, u = np.zeros(self.npacked), np.zeros(3 * self.nPoints)
        _check(load().hfls_recover(self._h, _ptr(f), _ptr(lam), _ptr(alpha), _ptr(x), _ptr(u)))
        return x, u

    def Rf(self):
        """Rf_d = R_d H_d per subdomain, packed local n_d x 6 column major."""
        out = np.zeros(6 * self.npacked)
        _check(load().hfls_mesh_rf(self._h, _ptr(out)))
        return out

    def counts(self):
        """The Cluster getters (Cluster.hpp:26-31)."""
        c = np.zeros(6, dtype=np.int64)
        _check(load().hfls_mesh_counts(self._h, _ptr(c)))
        return dict(zip(["nSubClst", "neqClst", "nLam_c", "nLam_f", "nRBM_f", "nRBM_c"], (int(x) for x in c)))

    def time_stream(self, which=0, reps=20):
        ms, nbytes = ctypes.c_double(), ctypes.c_double()
        _check(load().hfls_time_stream(self._h, int(which), int(reps), ctypes.byref(ms), ctypes.byref(nbytes)))
        return ms.value, nbytes.value

    def pcpg(self, eps=1e-8, max_iter=200, f_packed=None, want_lambda=True):
        u = np.zeros(3 * self.nPoints)
        lam = np.zeros(self.nlam) if want_lambda else None
        st = Stats()
        f = None if f_packed is None else _vec(f_packed, self.npacked, "f_packed")
        _check(load().hfls_solve(self._h, ctypes.c_double(eps), int(max_iter), _ptr(f), _ptr(u), _ptr(lam),
                                 ctypes.byref(st)))
        return {"u": u, "lambda": lam, "iterations": st.iterations, "converged": bool(st.converged),
                "residual": st.residual, "pg_rel": st.pg_rel, "solve_ms": st.solve_ms, "pcpg_ms": st.pcpg_ms}

    def traction(self, seed, case0, k):
        """k seeded smoothed tractions on the x = 1 face, packed local loads (k, npacked)."""
        out = np.zeros((k, self.npacked))
        _check(load().hfls_mesh_traction(self._h, ctypes.c_uint64(seed), int(case0), int(k), _ptr(out)))
        return out

    def apply_block(self, op, W, precision=64):
        """op 0 F, 1 M, 2 P applied to the rows of W (k, nlam) at precision 64 or 32."""
        W = _rows(W, self.nlam, "W")
        out = np.zeros_like(W)
        _check(load().hfls_apply_block(self._h, int(op), W.shape[0], int(precision), _ptr(W), _ptr(out)))
        return out

    def mult_Ff_block(self, W, precision=64):
        return self.apply_block(0, W, precision)

    def preconditioning_block(self, W, precision=64):
        return self.apply_block(1, W, precision)

    def projection_block(self, W):
        return self.apply_block(2, W, 64)

    def pcpg_block(self, F=None, k=None, eps=1e-8, max_iter=200, precision=32, want_lambda=False, out=None):
        """Block of load cases: F (k, npacked) packed local loads, or None for k
        copies of the body-force load.  Returns u (k, 3*nPoints), per-case
        iterations/status/residual and the block timings.  Status per case: 2
        converged, 3 iteration limit, 6 stopped on a non-finite residual.
        out: a C-contiguous float64 (k, 3*nPoints) array that receives u (for
        example a reused pinned_empty buffer); returned as "u"."""
        if F is not None:
            F = _rows(F, self.npacked, "F")
            k = F.shape[0]
        k = 1 if k is None else int(k)
        if out is None:
            u = np.zeros((k, 3 * self.nPoints))
        else:
            if (not isinstance(out, np.ndarray) or out.dtype != np.float64 or out.shape != (k, 3 * self.nPoints)
                    or not out.flags.c_contiguous or not out.flags.writeable):
                raise ValueError("out must be a writeable C-contiguous float64 array of shape (%d, %d)"
                                 % (k, 3 * self.nPoints))
            u = out
        lam = np.zeros((k, self.nlam)) if want_lambda else None
        cs = (CaseStats * k)()
        bs = BlockStats()
        _check(load().hfls_solve_block(self._h, ctypes.c_double(eps), int(max_iter), k, int(precision), _ptr(F),
                                       _ptr(u), _ptr(lam), cs, ctypes.byref(bs)))
        return {"u": u, "lambda": lam, "iterations": np.array([c.iterations for c in cs]),
                "status": np.array([c.status for c in cs]), "residual": np.array([c.residual for c in cs]),
                "converged": np.array([c.status == 2 for c in cs]), "solve_ms": bs.solve_ms, "init_ms": bs.init_ms,
                "pcpg_ms": bs.pcpg_ms, "recover_ms": bs.recover_ms, "sweeps": bs.sweeps,
                "cert_steps": bs.cert_steps}

    def pcpg_blocks(self, F, k=64, eps=1e-8, max_iter=200, precision=32):
        """F (nblocks*k, npacked): nblocks blocks of k load cases solved concurrently,
        one block context (own stream) each.  solve_ms is the device time span of all
        blocks; per-case results in F's order."""
        F = _rows(F, self.npacked, "F")
        if k < 1 or F.shape[0] % k:
            raise ValueError("the number of load cases must be a multiple of k")
        nb = F.shape[0] // k
        u = np.zeros((nb * k, 3 * self.nPoints))
        cs = (CaseStats * (nb * k))()
        bs = BlockStats()
        _check(load().hfls_solve_blocks(self._h, ctypes.c_double(eps), int(max_iter), nb, int(k), int(precision),
                                        _ptr(F), _ptr(u), cs, ctypes.byref(bs)))
        return {"u": u, "iterations": np.array([c.iterations for c in cs]),
                "status": np.array([c.status for c in cs]), "residual": np.array([c.residual for c in cs]),
                "solve_ms": bs.solve_ms, "init_ms": bs.init_ms, "pcpg_ms": bs.pcpg_ms, "recover_ms": bs.recover_ms,
                "sweeps": bs.sweeps, "cert_steps": bs.cert_steps, "blocks": nb}


class Server:
    """Request loop with continuous batching over device-cached meshes (one GPU).

    submit() queues k load cases for a mesh given by its options (packed local
    loads of shape (k, npacked), or seeded face tractions); step() runs one
    scheduling round (admission by HBM footprint, free slots take queued cases,
    one chunk of PCPG iterations per busy mesh, stopped cases recovered);
    result() returns the displacements and per-case stats of a finished request.
    """

    def __init__(self, device=0, hbm_budget_gb=200.0, slots=64, precision=32, eps=1e-8, max_iter=1000, chunk=8,
                 host_cache_gb=0.0):
        lib = load()
        self._h = ctypes.c_void_p()
        _check(lib.hfls_server_create(int(device), ctypes.c_double(hbm_budget_gb), int(slots), int(precision),
                                      ctypes.c_double(eps), int(max_iter), int(chunk), ctypes.byref(self._h)))
        self._k = {}
        if host_cache_gb:
            self.host_cache(host_cache_gb)

    def host_cache(self, gb):
        """Pinned host memory (GB, 0 = off) for the set-up state of evicted meshes: a
        later request for an evicted mesh restores it instead of setting it up."""
        _check(load().hfls_server_host_cache(self._h, ctypes.c_double(gb)))

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.hfls_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def sizes(self, **opts):
        """(npacked, 3*nPoints, estimated device bytes) of the mesh given by opts"""
        n, k, v = _opts(opts)
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _check(load().hfls_server_sizes(self._h, n, k, v, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def submit(self, opts, F=None, k=None, seed=0, case0=0):
        n, kk, v = _opts(opts)
        if F is not None:
            F = _rows(F, self.sizes(**opts)[0], "F")
            k = F.shape[0]
        k = 1 if k is None else int(k)
        rid = ctypes.c_int64()
        _check(load().hfls_server_submit(self._h, n, kk, v, k, _ptr(F), ctypes.c_uint64(seed), int(case0),
                                         ctypes.byref(rid)))
        self._k[rid.value] = (k, self.sizes(**opts)[1])
        return rid.value

    def step(self):
        busy = ctypes.c_int()
        _check(load().hfls_server_step(self._h, ctypes.byref(busy)))
        return bool(busy.value)

    def state(self, rid):
        s = ctypes.c_int()
        _check(load().hfls_server_state(self._h, ctypes.c_int64(rid), ctypes.byref(s)))
        return s.value

    def result(self, rid):
        """Displacements and per-case stats of a finished request (status 2
        converged, 3 iteration limit, 6 non-finite residual).  A request whose mesh
        set-up failed raises HflsError (its per-case status, 7, is on the error's
        ``status`` attribute)."""
        k, ng = self._k.pop(rid)
        u = np.zeros((k, ng))
        cs = (CaseStats * k)()
        lat = np.zeros(2)
        rc = load().hfls_server_result(self._h, ctypes.c_int64(rid), _ptr(u), cs, _ptr(lat))
        if rc != 0:
            err = HflsError(load().hfls_last_error().decode())
            err.status = np.array([c.status for c in cs])
            raise err
        return {"u": u, "iterations": np.array([c.iterations for c in cs]),
                "status": np.array([c.status for c in cs]), "residual": np.array([c.residual for c in cs]),
                "latency_ms": float(lat[0]), "queue_ms": float(lat[1])}

    def cancel(self, rid):
        """Drop a request: queued cases removed, running cases leave their slots."""
        _check(load().hfls_server_cancel(self._h, ctypes.c_int64(rid)))
        self._k.pop(rid, None)

    def stats(self):
        out = np.zeros(14)
        _check(load().hfls_server_stats(self._h, _ptr(out)))
        keys = ["meshes", "bytes", "setups", "evictions", "cases_done", "requests_done", "joins", "setup_ms",
                "failed_setups", "cancelled", "steps", "join_ms", "sweep_ms", "finish_ms"]
        st = dict(zip(keys, out.tolist()))
        c = np.zeros(7)
        _check(load().hfls_server_cache_stats(self._h, _ptr(c)))
        st.update(zip(["host_meshes", "host_bytes", "saves", "save_ms", "restores", "restore_ms", "host_drops"],
                      c.tolist()))
        return st

