"""Structured hex8 cube mesh (restates Mesh.cpp:69-289).  TEST INFRASTRUCTURE ONLY.

Extensions over the reference (documented, applied identically in the product):
  * clusters: ``Cx, Cy, Cz`` clusters per axis, each of ``Nx x Ny x Nz`` subdomains.
    Subdomain id = cluster_id * (Nx*Ny*Nz) + (i + j*Nx + k*Nx*Ny) with cluster_id =
    CI + CJ*Cx + CK*Cx*Cy.  With one cluster this is the reference's partition id
    ``II + JJ*Nx + KK*Nx*Ny`` (Mesh.cpp:271-288).
  * uneven slabs: per-axis element counts per subdomain slab (``slabs_x`` ...),
    uniform node spacing 1/NX as in the reference (Mesh.cpp:153-171).
  * boundary condition ``bc``: "symmetry" (reference, Mesh.cpp:221-238) or
    "clamp_x0" (all three DOFs fixed on x=0).
  * material ``material``: "reference" (E*ratio_mat where remainder(partId,2)==0,
    Mesh.cpp:281, Data.cpp:47-48), "uniform", or "lognormal" per subdomain.

  * ``metis=true`` with ``nparts=1`` (the reference's own main() map, Mesh.cpp:354-452):
    one subdomain over the whole cube; corner DOFs and material ids stay those of the
    geometric grid createMesh built (Mesh.cpp:413-421).

Parity: node/element/partition numbering, corner and Dirichlet DOF lists are
pinned by closed-form counts and hand-checked small cases in
tests/test_oracle_decomp.py (test_mesh_counts_and_partition, test_dirichlet_lists)
and by the committed integer fixtures of tests/golden; the single-subdomain map by a
global direct solve and by the geometric decomposition of the same cube
(tests/test_options.py).
"""
import math
import numpy as np

DEFAULTS = {
    "Nx": "2", "Ny": "2", "Nz": "2", "nx": "5", "ny": "5", "nz": "5",
    "Cx": "1", "Cy": "1", "Cz": "1",
    "young_modulus": "1000", "ratio_mat": "1e4", "poissons_ratio": "0.3",
    "material": "reference", "mat_sigma": "1.5", "mat_contrast": "1e3", "seed": "0",
    "bc": "symmetry", "load": "reference", "body_force": "-1.0", "density": "2.0",
    "typeBc": "ker", "Bc_fullRank": "true", "preconditioner": "Dirichlet_explicit",
    "eps_iter": "1e-6", "max_iter": "200", "Ac_extended_by_kerGc": "true",
    "slabs_x": "", "slabs_y": "", "slabs_z": "", "slab_choices": "",
    "metis": "false", "nparts": "1",
}


def opts(**kw):
    """Return an options2-like dict of strings with the defaults filled in."""
    o = dict(DEFAULTS)
    for k, v in kw.items():
        if isinstance(v, (list, tuple)):
            v = ",".join(str(int(x)) for x in v)
        o[k] = str(v)
    return o


def _slabs(o, axis):
    """Element counts of every subdomain slab along one axis (all clusters)."""
    C = int(o["C" + axis])
    N = int(o["N" + axis])
    s = o.get("slabs_" + axis, "")
    if s:
        v = [int(t) for t in s.split(",")]
        assert len(v) == C * N, "slabs_%s needs %d entries" % (axis, C * N)
        return v
    ch = o.get("slab_choices", "")
    if ch:
        # seeded draw from the list (uneven batching): choices[floor(u * len)], u from a
        # splitmix64 stream seeded with seed ^ (0xC2B2AE3D27D4EB4F * (axis index + 1))
        ch = [int(t) for t in ch.split(",")]
        ax = "xyz".index(axis) + 1
        u = uniforms((int(o["seed"]) ^ ((0xC2B2AE3D27D4EB4F * ax) & MASK64)) & MASK64, C * N)
        return [ch[min(int(x * len(ch)), len(ch) - 1)] for x in u]
    return [int(o["n" + axis])] * (C * N)


def lognormal_E(o, nsub):
    """Per-subdomain Young's modulus E_d = E*exp(clip(sigma*z_d, +-ln(contrast)/2)),
    z_d ~ N(0,1) from a splitmix64 stream seeded by ``seed`` (Box-Muller).  The
    product evaluates the same integer stream, so E_d agree to the last bit of the
    inputs to exp/log."""
    E = float(o["young_modulus"])
    sig = float(o["mat_sigma"])
    half = math.log(float(o["mat_contrast"])) / 2.0
    z = _normals(int(o["seed"]), nsub)
    out = np.empty(nsub)
    for d in range(nsub):
        t = min(max(sig * z[d], -half), half)
        out[d] = E * math.exp(t)
    return out


MASK64 = (1 << 64) - 1


def splitmix64(state):
    state = (state + 0x9E3779B97F4A7C15) & MASK64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return state, z ^ (z >> 31)


def uniforms(seed, n):
    """n doubles in (0,1): (top 53 bits + 0.5) / 2^53 of a splitmix64 stream."""
    st = seed & MASK64
    out = []
    for _ in range(n):
        st, z = splitmix64(st)
        out.append(((z >> 11) + 0.5) / 9007199254740992.0)
    return out


def _normals(seed, n):
    u = uniforms(seed, 2 * n)
    return [math.sqrt(-2.0 * math.log(u[2 * i])) * math.cos(2.0 * math.pi * u[2 * i + 1])
            for i in range(n)]


class Mesh:
    """Mesh::createMesh (Mesh.cpp:69-289) on the unit cube."""

    def __init__(self, o):
        self.o = o
        self.Cxyz = [int(o["Cx"]), int(o["Cy"]), int(o["Cz"])]
        self.Nxyz = [int(o["Nx"]), int(o["Ny"]), int(o["Nz"])]
        self.slabs = [_slabs(o, a) for a in "xyz"]
        # slab start indices along each axis
        self.starts = [np.concatenate([[0], np.cumsum(s)]).astype(np.int64) for s in self.slabs]
        NX, NY, NZ = (int(st[-1]) for st in self.starts)
        self.NE = (NX, NY, NZ)
        self.nsub = int(np.prod(self.Cxyz) * np.prod(self.Nxyz))
        self.nsub_cl = int(np.prod(self.Nxyz))
        self.ncl = int(np.prod(self.Cxyz))
        nx1, ny1, nz1 = NX + 1, NY + 1, NZ + 1
        self.nPoints = nx1 * ny1 * nz1
        # points (Mesh.cpp:152-171): x = (ii*dx - 0.5) + 0.5 etc.; no curvature
        # because radius=1050 > 100*radius_min (Mesh.cpp:133-147).
        dx = 1.0 / NX
        dy = 1.0 / NY
        dz = 1.0 / NZ
        kk, jj, ii = np.meshgrid(np.arange(nz1), np.arange(ny1), np.arange(nx1), indexing="ij")
        x = (ii.ravel() * dx - 0.5) + 0.5
        y = (jj.ravel() * dy - 0.5) + 0.5
        z = (kk.ravel() * dz) + 0.0
        self.points = np.stack([x, y, z], axis=1)
        self.grid = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1).astype(np.int64)
        # elements (Mesh.cpp:243-264)
        ek, ej, ei = np.meshgrid(np.arange(NZ), np.arange(NY), np.arange(NX), indexing="ij")
        ei, ej, ek = ei.ravel(), ej.ravel(), ek.ravel()
        nxy = nx1 * ny1
        f0 = ei + ej * nx1
        f1 = ei + 1 + ej * nx1
        f2 = ei + 1 + (ej + 1) * nx1
        f3 = ei + (ej + 1) * nx1
        first = np.stack([f0, f1, f2, f3], axis=1)
        self.elements = np.concatenate([first + ek[:, None] * nxy, first + (ek[:, None] + 1) * nxy],
                                       axis=1).astype(np.int64)
        # partition ids (Mesh.cpp:271-288), generalised to clusters and slabs
        I = np.searchsorted(self.starts[0], ei, side="right") - 1
        J = np.searchsorted(self.starts[1], ej, side="right") - 1
        K = np.searchsorted(self.starts[2], ek, side="right") - 1
        self.part = self.sub_id(I, J, K)
        # material id: (int)remainder(partId, 2) with partId the partition counter
        self.material_id = np.array([int(math.remainder(p, 2)) for p in range(self.nsub)])
        self.corner_dofs = self._corners()
        self.dirichlet_dofs = self._dirichlet(o["bc"])
        # geometric partition of every element; it keeps the material ids when metis
        # merges the partition (Mesh.cpp:271-288, 425-427)
        self.geo_part = self.part.copy()
        self.geo_material_id = self.material_id.copy()
        if o.get("metis", "false") == "true":
            # Mesh::ddm_metis with nparts <= 1 puts every element in partition 0
            # without calling METIS (Mesh.cpp:413-421): one subdomain, one cluster
            if int(o.get("nparts", "1")) > 1:
                raise ValueError("metis=true with nparts > 1 needs METIS")
            if self.ncl != 1:
                raise ValueError("metis=true needs one cluster")
            self.Nxyz = [1, 1, 1]
            self.slabs = [[NX], [NY], [NZ]]
            self.starts = [np.array([0, n], dtype=np.int64) for n in (NX, NY, NZ)]
            self.nsub = self.nsub_cl = 1
            self.part = np.zeros_like(self.part)
            self.material_id = np.array([0])

    def sub_id(self, I, J, K):
        Nx, Ny, Nz = self.Nxyz
        Cx, Cy, _ = self.Cxyz
        cl = (I // Nx) + (J // Ny) * Cx + (K // Nz) * Cx * Cy
        loc = (I % Nx) + (J % Ny) * Nx + (K % Nz) * Nx * Ny
        return (cl * self.nsub_cl + loc).astype(np.int64)

    def cluster_of(self, d):
        return d // self.nsub_cl

    def _corners(self):
        """cornerDOFs (Mesh.cpp:175-208): subdomain-grid vertices without the 8 cube
        corners, k outer, j, i inner."""
        NX, NY, NZ = self.NE
        nx1, nxy1 = NX + 1, (NX + 1) * (NY + 1)
        cube = {0, NX, NY * nx1, NY * nx1 + NX}
        cube |= {c + NZ * nxy1 for c in list(cube)}
        out = []
        for k in self.starts[2]:
            for j in self.starts[1]:
                for i in self.starts[0]:
                    n = int(i + nx1 * j + nxy1 * k)
                    if n not in cube:
                        out += [3 * n, 3 * n + 1, 3 * n + 2]
        return np.array(out, dtype=np.int64)

    def _dirichlet(self, bc):
        """DirichletDOFs (Mesh.cpp:221-238): points ascending, components x,y,z."""
        eps0 = 1e-8
        p = self.points
        out = []
        if bc == "symmetry":
            for i in range(self.nPoints):
                if abs(p[i, 0]) < eps0:
                    out.append(3 * i)
                if abs(p[i, 1]) < eps0:
                    out.append(3 * i + 1)
                if abs(p[i, 2]) < eps0:
                    out.append(3 * i + 2)
        elif bc == "clamp_x0":
            for i in np.nonzero(np.abs(p[:, 0]) < eps0)[0]:
                out += [3 * i, 3 * i + 1, 3 * i + 2]
        else:
            raise ValueError("unknown bc " + bc)
        return np.array(out, dtype=np.int64)

    def young_per_element(self):
        """E per element: the reference material follows the geometric partition's
        material id (Data.cpp:47-48), the other modes the subdomain's E."""
        if self.o["material"] == "reference":
            E = float(self.o["young_modulus"])
            r = float(self.o["ratio_mat"])
            return np.where(self.geo_material_id[self.geo_part] == 0, E * r, E)
        return self.young_per_sub()[self.part]

    def young_per_sub(self):
        """E per subdomain (Data.cpp:36,47-48 for the reference mode)."""
        o = self.o
        E = float(o["young_modulus"])
        mode = o["material"]
        if mode == "reference":
            r = float(o["ratio_mat"])
            return np.array([E * r if m == 0 else E for m in self.material_id])
        if mode == "uniform":
            return np.full(self.nsub, E)
        if mode == "lognormal":
            return lognormal_E(o, self.nsub)
        raise ValueError("unknown material " + mode)
