"""The C ABI library loads and exports every function include/fluidsim_hip.h declares (CPU only: no
compute call is made here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fluidsim_hip.h")
LIB = os.path.join(ROOT, "fluidsim_amd", "libfluidsim_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fs_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("fs_context_create", "fs_grid_create", "fs_upload", "fs_download", "fs_step", "fs_project",
                     "fs_max_divergence", "fs_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "libfluidsim_hip.so is not built (python -m fluidsim_amd.build)"
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"missing exports: {missing}"


def test_python_binding_covers_the_header():
    from fluidsim_amd import _abi

    bound = set(_abi.SIGNATURES) | {"fs_last_error", "fs_version"}
    assert set(declared_functions()) <= bound


def test_version_and_error_string_need_no_device():
    lib = ctypes.CDLL(LIB)
    lib.fs_version.restype = ctypes.c_char_p
    lib.fs_last_error.restype = ctypes.c_char_p
    assert b"gfx950" in lib.fs_version()
    assert lib.fs_last_error() is not None


def test_invalid_arguments_are_rejected_before_any_device_work():
    from fluidsim_amd import _abi

    L = _abi.lib()
    assert L.fs_grid_create(None, 4, 4, 1.0, None) == _abi.FS_E_INVALID_ARGUMENT
    assert b"null" in L.fs_last_error()
    assert L.fs_project(None, 0.1, 10, 1e-6, None) == _abi.FS_E_INVALID_ARGUMENT
    with pytest.raises(ValueError):
        _abi.check(_abi.FS_E_INVALID_ARGUMENT, "probe")


def test_server_and_estimate_validation_need_no_device():
    import ctypes as C

    from fluidsim_amd import _abi

    L = _abi.lib()
    cfg = _abi.FsServerConfig(0, 0, 1 << 30)
    h = C.c_void_p()
    assert L.fs_server_create(C.byref(cfg), C.byref(h)) == _abi.FS_E_INVALID_ARGUMENT
    assert b"lanes" in L.fs_last_error()
    b = C.c_size_t()
    assert L.fs_grid_estimate_bytes(0, 8, 1, 1, 0, C.byref(b)) == _abi.FS_E_INVALID_ARGUMENT
    sizes = []
    for features in (0, 1, 2, 4, 7):
        assert L.fs_grid_estimate_bytes(256, 128, 1, 1, features, C.byref(b)) == _abi.FS_OK
        sizes.append(b.value)
    assert sizes[0] < min(sizes[1:4]) and sizes[4] > max(sizes[1:4])
    assert L.fs_grid_estimate_bytes(512, 256, 1, 1, 0, C.byref(b)) == _abi.FS_OK and b.value > 3 * sizes[0]
