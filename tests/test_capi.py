"""The C-ABI library loads and exports every function include/awk_hip.h declares
(no device calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "awk_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(awk_[a-z0-9_]+)\s*\(", src)))


def test_exports_every_declared_symbol():
    from atlaswerks_amd import _lib
    lib = _lib.load()
    names = declared()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_error_path_without_device():
    from atlaswerks_amd import _lib
    lib = _lib.load()
    assert lib.awk_version() == 100
    h = ctypes.c_void_p()
    # invalid sizes are rejected before any device call
    rc = lib.awk_fluid_kernel_create(ctypes.byref(h), 0, 4, 4, 1.0, 1.0, 1.0, 0.01, 0.01, 0.001, 1.0, 0)
    assert rc == _lib.AWK_ERR_INVALID
    assert b"sizes" in lib.awk_last_error()
    rc = lib.awk_tune_set(b"nonsense", 1)
    assert rc == _lib.AWK_ERR_INVALID
