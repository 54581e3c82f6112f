"""The allocation-free projection (hfls_project_noalloc) through the C ABI against
hfls_project (Projection, Cluster.cpp:1314-1329) on the same inputs: bit-identical
P x and alpha, on two calls in a row with the second on a smaller mesh (the shared
scratch array keeps the first call's contents), both on the library's stream and
enqueued on a caller stream.  The caller's device buffers and stream come from the
HIP runtime the library itself is linked against (hipMalloc / hipStreamCreate
through ctypes in the test, never inside the entry point).
"""
import ctypes

import numpy as np
import pytest

import hfls_amd

pytestmark = pytest.mark.gpu

BIG = dict(Cx=2, Cy=1, Cz=1, Nx=3, Ny=2, Nz=2, nx=3, ny=3, nz=3, bc="clamp_x0", material="lognormal",
           young_modulus="2.1e5", seed="4", load="body")
SMALL = dict(Nx=2, Ny=2, Nz=1, nx=3, ny=2, nz=4, bc="clamp_x0", material="uniform", young_modulus="2.1e5",
             load="body")
H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def hip_runtime():
    """The libamdhip64 already loaded by libhfls.so (same runtime instance)."""
    hfls_amd.load()
    with open("/proc/self/maps") as f:
        paths = {line.split()[-1] for line in f if "libamdhip64" in line}
    assert paths, "libamdhip64 is not loaded"
    return ctypes.CDLL(sorted(paths)[0])


class DeviceBuffer:
    def __init__(self, hip, n):
        self.hip, self.n = hip, n
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(8 * n)) == 0

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(8 * self.n), H2D) == 0

    def download(self):
        a = np.empty(self.n)
        assert self.hip.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), self.p, ctypes.c_size_t(8 * self.n), D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


def noalloc(hip, mesh, x, stream):
    lib = hfls_amd.load()
    xd, Pd, ad = DeviceBuffer(hip, mesh.nlam), DeviceBuffer(hip, mesh.nlam), DeviceBuffer(hip, mesh.ncol)
    try:
        xd.upload(x)
        Pd.upload(np.full(mesh.nlam, np.nan))
        ad.upload(np.full(mesh.ncol, np.nan))
        assert hip.hipDeviceSynchronize() == 0
        rc = lib.hfls_project_noalloc(mesh._h, xd.p, Pd.p, ad.p, stream)
        if rc != 0:
            raise hfls_amd.HflsError(lib.hfls_last_error().decode())
        if stream is not None:
            assert hip.hipStreamSynchronize(stream) == 0
        return Pd.download(), ad.download()
    finally:
        for b in (xd, Pd, ad):
            b.free()


@pytest.mark.parametrize("use_stream", [False, True])
def test_noalloc_projection_matches(use_stream):
    hip = hip_runtime()
    stream = None
    if use_stream:
        stream = ctypes.c_void_p()
        assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    rng = np.random.default_rng(17)
    try:
        for kw in (BIG, SMALL):  # the second call on a smaller mesh
            print("mesh", kw, "caller stream" if use_stream else "library stream", flush=True)
            m = hfls_amd.Mesh(**kw)
            try:
                x = rng.standard_normal(m.nlam)
                Px, alpha = m.projection(x)
                Pn, an = noalloc(hip, m, x, stream)
                assert np.array_equal(Pn, Px) and np.array_equal(an, alpha)
                # again on the same mesh (the scratch holds this mesh's partials now)
                x2 = rng.standard_normal(m.nlam)
                P2, a2 = m.projection(x2)
                Pn2, an2 = noalloc(hip, m, x2, stream)
                assert np.array_equal(Pn2, P2) and np.array_equal(an2, a2)
            finally:
                m.close()
    finally:
        if stream is not None:
            hip.hipStreamDestroy(stream)


def test_noalloc_rejects_null_buffers():
    m = hfls_amd.Mesh(**SMALL)
    try:
        lib = hfls_amd.load()
        assert lib.hfls_project_noalloc(m._h, None, None, None, None) != 0
        assert b"null argument" in lib.hfls_last_error()
    finally:
        m.close()
