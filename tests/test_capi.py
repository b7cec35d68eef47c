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


def test_knob_ranges_without_device():
    """Every knob accepts exactly its documented values and reads back what was set."""
    from atlaswerks_amd import _lib
    lib = _lib.load()
    table = {"apply_vpt": (1, 2, 4), "compose_vpt": (1, 2, 4), "xinv_unroll": (1, 2, 4),
             "xinv_pairs": (0, 1), "splat_tile": (0, 1), "unet_lds": (0, 1), "unet_pipe": (0, 1, 2),
             "unet_cc": (0, 8, 16, 32), "atlas_mean": (0, 1, 2, 4), "xinv_band": (0, 2, 4, 8, 16, 32), "xfwd_band": (0, 2, 4, 8, 16, 32), "yfwd_rev": (0, 1), "xinv_stage": (0, 1, 2), "fft_order": (0, 1),
             "x_rt": (0, 2, 4, 6, 8), "y_tx": (0, 1, 2, 4, 8, 16), "z_tx": (0, 1, 2, 4, 8)}
    for name, ok in table.items():
        v = ctypes.c_int()
        assert lib.awk_tune_get(name.encode(), ctypes.byref(v)) == 0, name
        saved = v.value
        assert saved in ok, (name, saved)
        for x in range(-1, 34):
            rc = lib.awk_tune_set(name.encode(), x)
            assert (rc == 0) == (x in ok), (name, x, rc)
            if rc == 0:
                assert lib.awk_tune_get(name.encode(), ctypes.byref(v)) == 0 and v.value == x
        assert lib.awk_tune_set(name.encode(), saved) == 0
