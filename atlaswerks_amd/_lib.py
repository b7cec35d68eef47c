"""ctypes binding of libawk_hip.so (the C-ABI declared in include/awk_hip.h).

The product path is the HIP library only: if ``libawk_hip.so`` is missing or
fails to load this module raises; there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libawk_hip.so")

AWK_OK, AWK_ERR_INVALID, AWK_ERR_HIP, AWK_ERR_ALLOC, AWK_ERR_UNSUPPORTED, AWK_ERR_NUMERIC = range(6)
FLAG_NAN_ENERGY, FLAG_ENERGY_INCREASE, FLAG_BAD_STEP = 1, 2, 4
BG_ID, BG_ZERO, BG_PARTIAL_ID, BG_PARTIAL_ZERO, BG_CLAMP, BG_WRAP, BG_VAL = range(7)
FK_EIGEN, FK_DIVFREE = 1, 2
SPLAT_FIXED64, SPLAT_FIXED20, SPLAT_FLOAT = range(3)

# (name, argtypes); every function returns awk_status (int) unless listed in _RESTYPE.
_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_P3 = ctypes.c_void_p * 3
SIGNATURES = {
    "awk_last_error": [],
    "awk_version": [],
    "awk_tune_set": [ctypes.c_char_p, _I],
    "awk_tune_get": [ctypes.c_char_p, ctypes.POINTER(_I)],
    "awk_fluid_kernel_create": [ctypes.POINTER(_P), _I, _I, _I, _F, _F, _F, _F, _F, _F, _F, _I],
    "awk_fluid_kernel_set_params": [_P, _F, _F, _F, _F, _I],
    "awk_fluid_kernel_destroy": [_P],
    "awk_fluid_kernel_scratch_bytes": [_P, ctypes.POINTER(ctypes.c_ulonglong)],
    "awk_fluid_kernel_apply": [_P, _P, _P, _P, _I, _P],
    "awk_hfield_set_to_identity": [_P, _P, _P, _I, _I, _I, _P],
    "awk_hfield_velocity_to_h": [_P, _P, _P, _I, _I, _I, _F, _F, _F, _P],
    "awk_hfield_h_to_velocity": [_P, _P, _P, _I, _I, _I, _F, _F, _F, _P],
    "awk_hfield_inverse_zeroth_order": [_P, _P, _P, _P, _P, _P, _I, _I, _I, _P],
    "awk_hfield_apply": [_P, _P, _P, _P, _P, _I, _I, _I, _I, _F, _P],
    "awk_hfield_apply_grid": [_P, _P, _I, _I, _I, _P, _P, _P, _I, _I, _I, _I, _F, _P],
    "awk_hfield_init_affine": [_P, _P, _P, _I, _I, _I, _P, _P],
    "awk_hfield_compose": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _P],
    "awk_hfield_compose_hv": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _F, _F, _F, _F, _I,
                              _I, _P],
    "awk_hfield_compose_vh": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _F, _F, _F, _I, _I,
                              _P],
    "awk_gradient": [_P, _P, _P, _P, _I, _I, _I, _F, _F, _F, _I, _P],
    "awk_splat": [_P, _P, _P, _P, _P, _I, _I, _I, _I, _P],
    "awk_splat_noalloc": [_P, _P, _P, _P, _P, _I, _I, _I, _I, _P],
    "awk_splat_sorted": [_P, _P, _P, _P, _P, _I, _I, _I, _I, _P, _P, _P],
    "awk_greedy_create": [ctypes.POINTER(_P), _I, _I, _I, _F, _F, _F, _F, _F, _F, _F],
    "awk_greedy_destroy": [_P],
    "awk_gaussian_filter": [_P, _P, _I, _I, _I, _F, _F, _F, _I, _I, _I, _P],
    "awk_gaussian_downsample": [_P, _I, _I, _I, _P, _I, _I, _I, _P],
    "awk_hfield_resample": [_P, _P, _P, _I, _I, _I, _P, _P, _P, _I, _I, _I, _I, _I, _P],
    "awk_hfield_add_affine": [_P, _P, _P, _I, _I, _I, _P, _F, _P],
    "awk_hfield_affine_displace": [_P, _P, _P, _P, _P, _P, _I, _I, _I, _P, _P],
    "awk_atlas_create": [ctypes.POINTER(_P), _I, _I, _I, _F, _F, _F, _F, _F, _F, _F, _I],
    "awk_atlas_destroy": [_P],
    "awk_atlas_set_subject": [_P, _I, _P, _I, _I, _I, _F, ctypes.POINTER(_P3), _P],
    "awk_atlas_set_subject_inverse": [_P, _I, _P, _P],
    "awk_atlas_get_hinv": [_P, _I, _P, _P],
    "awk_atlas_get_hfield": [_P, _I, ctypes.POINTER(_P3), _P],
    "awk_atlas_update_step_next": [_P],
    "awk_atlas_set_graph": [_P, _I],
    "awk_atlas_set_batch": [_P, _I],
    "awk_atlas_iterate": [_P, _I, _P, _P],
    "awk_atlas_mean": [_P, _P, _P],
    "awk_atlas_deformed_image": [_P, _I, _P, _P],
    "awk_unet_weight_count": [],
    "awk_unet_create": [ctypes.POINTER(_P), _I, _I, _I, _P, ctypes.c_size_t],
    "awk_unet_create_batch": [ctypes.POINTER(_P), _I, _I, _I, _I, _P, ctypes.c_size_t],
    "awk_unet_destroy": [_P],
    "awk_unet_forward": [_P, _P, _P, ctypes.POINTER(_P3), _P],
    "awk_unet_forward_batch": [_P, _I, _P, _P, _P, _P],
    "awk_unet_mfma_probe": [_P, _P],
    "awk_unet_trainer_create": [ctypes.POINTER(_P), _P, _P, ctypes.c_size_t],
    "awk_unet_trainer_destroy": [_P],
    "awk_unet_trainer_step": [_P, _P, _P, ctypes.POINTER(_P3), _F, _P, _P],
    "awk_unet_trainer_gradients": [_P, _P, _P, ctypes.POINTER(_P3), _P, ctypes.POINTER(ctypes.c_double)],
    "awk_unet_trainer_get_weights": [_P, _P, ctypes.c_size_t],
    "awk_ldmm_create": [ctypes.POINTER(_P), _I, _I, _I, _F, _F, _F, _F, _F, _F, _I, _F, _F],
    "awk_ldmm_destroy": [_P],
    "awk_ldmm_set_gradient": [_P, _I],
    "awk_ldmm_set_splat": [_P, _I],
    "awk_ldmm_update_step_next": [_P],
    "awk_ldmm_set_step": [_P, _F, _P],
    "awk_ldmm_get_step": [_P, ctypes.POINTER(_F), _P],
    "awk_ldmm_iterate": [_P, _P, _P, _P, _P, _I, _P],
    "awk_ldmm_alphas": [_P, _P, _P, _P, _P, _P],
    "awk_ldmm_deformed_image": [_P, _I, _P, _P],
    "awk_ldmm_atlas_create": [ctypes.POINTER(_P), _I, _I, _I, _F, _F, _F, _F, _F, _F, _I, _F, _F, _I, _I],
    "awk_ldmm_atlas_destroy": [_P],
    "awk_ldmm_atlas_set_subject": [_P, _I, _P, _F, _P, _P],
    "awk_ldmm_atlas_set_gradient": [_P, _I],
    "awk_ldmm_atlas_set_splat": [_P, _I],
    "awk_ldmm_atlas_update_step_next": [_P],
    "awk_ldmm_atlas_set_step": [_P, _F, _P],
    "awk_ldmm_atlas_get_step": [_P, _I, ctypes.POINTER(_F), _P],
    "awk_ldmm_atlas_iterate": [_P, _I, _P, _I, _P],
    "awk_ldmm_atlas_mean": [_P, _P, _P],
    "awk_ldmm_atlas_deformed_image": [_P, _I, _I, _P, _P],
    "awk_greedy_update_step_next": [_P],
    "awk_greedy_set_step": [_P, _F, _P],
    "awk_greedy_get_step": [_P, ctypes.POINTER(_F), _P],
    "awk_greedy_set_auto_step_reduce": [_P, _I],
    "awk_greedy_set_profiling": [_P, _I],
    "awk_greedy_get_profile": [_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_I), ctypes.POINTER(_I)],
    "awk_greedy_get_flags": [_P, ctypes.POINTER(ctypes.c_uint), _I, _P],
    "awk_vf_aos_to_soa": [_P, _P, _P, _P, ctypes.c_longlong, _P],
    "awk_vf_soa_to_aos": [_P, _P, _P, _P, ctypes.c_longlong, _P],
    "awk_atlas_get_flags": [_P, ctypes.POINTER(ctypes.c_uint), _I, _P],
    "awk_ldmm_get_flags": [_P, ctypes.POINTER(ctypes.c_uint), _I, _P],
    "awk_ldmm_atlas_get_flags": [_P, ctypes.POINTER(ctypes.c_uint), _I, _P],
    "awk_greedy_iterate": [_P, _P, _P, ctypes.POINTER(_P3), ctypes.POINTER(_P3),
                           ctypes.POINTER(_P3), ctypes.POINTER(_P3), _P, _P],
    "awk_greedy_iterate_n": [_P, _P, _P, ctypes.POINTER(_P3), ctypes.POINTER(_P3), _I, _P, _P],
}
_RESTYPE = {"awk_last_error": ctypes.c_char_p, "awk_version": ctypes.c_int,
            "awk_unet_weight_count": ctypes.c_size_t}

_lib = None


class AwkError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("awk status %d: %s" % (code, msg))
        self.code = code


class NumericalError(AwkError, FloatingPointError):
    """AWK_ERR_NUMERIC: a NaN energy or an invalid step size (the reference's
    "NaN encountered" / "max displacement is NaN" exceptions)."""


def load():
    """Load libawk_hip.so, raising if it is absent (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libawk_hip.so not built (run __graft_entry__.build()): %s" % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, ctypes.c_int)
        _lib = lib
    return _lib


def check(code):
    if code != AWK_OK:
        cls = NumericalError if code == AWK_ERR_NUMERIC else AwkError
        raise cls(code, load().awk_last_error().decode())
    return code


def get_flags(name, handle, clear=True, stream=None):
    """(flags, status) of an awk_*_get_flags call, without raising on
    AWK_ERR_NUMERIC; waits for the stream."""
    from .device import stream_handle
    f = ctypes.c_uint(0)
    code = getattr(load(), name)(handle, ctypes.byref(f), int(bool(clear)), stream_handle(stream))
    if code not in (AWK_OK, AWK_ERR_NUMERIC):
        check(code)
    return f.value, code


def call(name, *args):
    return check(getattr(load(), name)(*args))
