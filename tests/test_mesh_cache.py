"""Saved-mesh files (hfls_mesh_save / hfls_mesh_restore): the checks that run before
any device work, on the CPU.  A file that is missing, is not a saved mesh, has
another format version or ends inside its header is rejected with a message, and
the null-handle path of the save entry point returns an error instead of crashing.
The GPU side (bit-identical restores, checks of sizes and indices against the
mesh the options describe) is in test_gpu_cache.py."""
import ctypes
import struct

import pytest

import hfls_amd


def _restore_error(path):
    with pytest.raises(hfls_amd.HflsError) as e:
        hfls_amd.Mesh.restore(str(path))
    return str(e.value)


def _header(opts, version=1):
    b = b"HFLSMESH" + struct.pack("<II", version, len(opts))
    for k, v in sorted(opts.items()):
        for s in (k.encode(), v.encode()):
            b += struct.pack("<I", len(s)) + s
    return b


def test_missing_file(tmp_path):
    assert "cannot open" in _restore_error(tmp_path / "absent.mesh")


def test_not_a_saved_mesh(tmp_path):
    p = tmp_path / "x.mesh"
    p.write_bytes(b"# vtk DataFile Version 2.0\n" + b"\0" * 64)
    assert "bad magic" in _restore_error(p)


def test_other_format_version(tmp_path):
    p = tmp_path / "x.mesh"
    p.write_bytes(_header({"Nx": "2"}, version=99))
    assert "format version 99" in _restore_error(p)


def test_truncated_header(tmp_path):
    full = _header(dict(hfls_amd.DEFAULTS))
    p = tmp_path / "x.mesh"
    for cut in (4, 12, 20, len(full) // 2):
        p.write_bytes(full[:cut])
        assert "truncated" in _restore_error(p)


def test_oversized_option_count(tmp_path):
    p = tmp_path / "x.mesh"
    p.write_bytes(b"HFLSMESH" + struct.pack("<II", 1, 1 << 20))
    assert "too many options" in _restore_error(p)


def test_save_with_null_handle(tmp_path):
    lib = hfls_amd.load()
    assert lib.hfls_mesh_save(None, str(tmp_path / "y.mesh").encode()) != 0
    assert b"null mesh handle" in lib.hfls_last_error()
    out = ctypes.c_void_p()
    assert lib.hfls_mesh_restore(0, None, ctypes.byref(out), None) != 0
    assert b"null argument" in lib.hfls_last_error()
