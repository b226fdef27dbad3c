"""ctypes binding of libawig.so (the C ABI declared in include/awig.h).

The library is built in-tree by ``__graft_entry__.build()`` (or ``python -m
accuratewigner_amd.build``).  Importing this module does not need a GPU; calling into the
library does.  A missing library raises immediately: there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libawig.so")

AW_OK = 0
AW_PREC_F64 = 0
AW_PREC_F32 = 1
AW_PREC_C64 = 2   # complex64 states, FP64 arithmetic
AW_FORWARD = -1
AW_INVERSE = 1


# Exception classes mirror the reference FFT wrapper's error map (cufft.py:31-104).
class awigError(Exception):
    """libawig error"""


class awigInvalidPlan(awigError):
    """libawig was passed an invalid plan handle."""


class awigAllocFailed(awigError):
    """libawig failed to allocate GPU memory."""


class awigInvalidType(awigError):
    """The user requested an unsupported type."""


class awigInvalidValue(awigError):
    """The user specified a bad argument or memory pointer."""


class awigInternalError(awigError):
    """Internal error."""


class awigExecFailed(awigError):
    """A kernel failed to launch or execute on the GPU."""


class awigSetupFailed(awigError):
    """The library failed to initialize."""


class awigInvalidSize(awigError):
    """The user specified an unsupported transform size."""


class awigRTCError(awigError):
    """Runtime compilation (hipRTC) of a user expression failed."""


class awigHIPError(awigError):
    """A HIP runtime call failed."""


_EXCEPTIONS = {
    1: awigInvalidPlan, 2: awigAllocFailed, 3: awigInvalidType, 4: awigInvalidValue,
    5: awigInternalError, 6: awigExecFailed, 7: awigSetupFailed, 8: awigInvalidSize,
    9: awigRTCError, 10: awigHIPError,
}


class aw_pass(ctypes.Structure):
    """Mirror of ``struct aw_pass`` in include/awig.h."""
    _fields_ = [
        ("axis", ctypes.c_int), ("dir_a", ctypes.c_int), ("dir_b", ctypes.c_int),
        ("sign_in", ctypes.c_int), ("sign_out", ctypes.c_int), ("out_mode", ctypes.c_int),
        ("line_fac", ctypes.c_void_p), ("line_fac_bstride", ctypes.c_longlong),
        ("elem_fac", ctypes.c_void_p), ("elem_fac_bstride", ctypes.c_longlong),
        ("mask", ctypes.c_void_p), ("mask_bstride", ctypes.c_longlong),
        ("chirp", ctypes.c_double), ("lstep", ctypes.c_double), ("lcenter", ctypes.c_double),
        ("estep", ctypes.c_double), ("ecenter", ctypes.c_double),
        ("lfftorder", ctypes.c_int), ("efftorder", ctypes.c_int),
        ("scale", ctypes.c_double), ("dscale", ctypes.c_void_p), ("out", ctypes.c_void_p),
        ("ctab", ctypes.c_void_p), ("ctab_bstride", ctypes.c_longlong), ("chirp_out", ctypes.c_double),
        ("obs_weights", ctypes.c_void_p), ("obs_nw", ctypes.c_int), ("obs_partials", ctypes.c_void_p),
        ("obs_mode", ctypes.c_int), ("obs_norm2", ctypes.c_int), ("obs_norm2_weights", ctypes.c_void_p),
        ("in_scaled_out", ctypes.c_void_p),
        ("obs_line_fac", ctypes.c_void_p), ("obs_elem_fac", ctypes.c_void_p), ("obs_diag_nw", ctypes.c_int),
    ]


class aw_steps(ctypes.Structure):
    """Mirror of ``struct aw_steps`` in include/awig.h."""
    _fields_ = [
        ("nsteps", ctypes.c_int), ("const_steps", ctypes.c_int), ("batch", ctypes.c_int),
        ("col_line", ctypes.c_void_p), ("col_elem", ctypes.c_void_p),
        ("row_line", ctypes.c_void_p), ("row_elem", ctypes.c_void_p),
        ("col_line_bstride", ctypes.c_longlong), ("col_elem_bstride", ctypes.c_longlong),
        ("row_line_bstride", ctypes.c_longlong), ("row_elem_bstride", ctypes.c_longlong),
        ("mask_open_close", ctypes.c_void_p), ("mask_mid", ctypes.c_void_p), ("mask_row", ctypes.c_void_p),
        ("mask_bstride", ctypes.c_longlong),
        ("dscale", ctypes.c_void_p),
    ]


_vp, _i, _ll, _d, _sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_double, ctypes.c_size_t

# name -> argtypes (all functions return int status except noted)
_SIGNATURES = {
    "aw_version": [],
    "aw_plan_create": [ctypes.POINTER(_vp), _i, _i, _i, _i],
    "aw_plan_destroy": [_vp],
    "aw_fft_axis": [_vp, _vp, _vp, _i, _i, _vp],
    "aw_fft_r2c": [_vp, _vp, _vp, _i, _vp],
    "aw_fft_c2r": [_vp, _vp, _vp, _i, _vp],
    "aw_pass_run": [_vp, _vp, ctypes.POINTER(aw_pass), _vp],
    "aw_row_obs_blocks": [_vp, ctypes.POINTER(_i)],
    "aw_pass_obs_blocks": [_vp, _i, ctypes.POINTER(_i)],
    "aw_step_sequence": [_vp, _vp, ctypes.POINTER(aw_steps), _vp],
    "aw_step_group": [_vp, _i, _i, ctypes.POINTER(_i)],
    "aw_make_factors": [_vp, _vp, _ll, _i, _i, _i, _vp, _ll, _d, _vp, _ll, _d, _d, _vp, _ll, _i,
                        _d, _i, _vp],
    "aw_reduce_stats": [_vp, _vp, _vp, _vp],
    "aw_reduce_vec": [_i, _vp, _ll, _i, _vp, _vp],
    "aw_sum_partials": [_vp, _i, _i, _i, _vp, _vp],
    "aw_finish_sums": [_vp, _i, _i, _i, _vp, _d, _vp, _vp],
    "aw_pair_dot": [_vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "aw_cyclic_diag_sums": [_vp, _vp, _vp, _vp],
    "aw_weighted_sums": [_vp, _vp, _vp, _i, _i, _vp, _d, _vp, _vp],
    "aw_weighted_partials": [_vp, _vp, _vp, _i, _i, _vp, ctypes.POINTER(_i), _vp],
    "aw_weighted_sums_noalloc": [_i, _i, _i, _i, _vp, _vp, _i, _i, _vp, _d, _vp, _vp],
    "aw_cyclic_diag_obs": [_vp, _vp, _vp, _i, _vp, _i, _vp, _ll, _d, _vp, _vp],
    "aw_scale_from_stats": [_vp, _i, _i, _i, _d, _vp, _i, _vp],
    "aw_scale": [_i, _vp, _ll, _i, _vp, _d, _vp],
    "aw_convert": [_vp, _i, _vp, _i, _ll, _vp],
    "aw_module_compile": [ctypes.POINTER(_vp), ctypes.c_char_p, ctypes.c_char_p,
                          ctypes.c_char_p, _sz],
    "aw_module_destroy": [_vp],
    "aw_rtc_check": [ctypes.c_char_p, ctypes.c_char_p, _sz, ctypes.POINTER(_sz)],
    "aw_module_launch": [_vp, ctypes.c_char_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
                         ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, _vp,
                         ctypes.POINTER(_vp)],
    "aw_project_out": [_i, _vp, _vp, _i, _i, _vp],
    "aw_mem_info": [ctypes.POINTER(_sz), ctypes.POINTER(_sz)],
    "aw_sync": [_vp],
}

_lib = None


def load():
    """Load libawig.so (raises if it is not built: the product has no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise awigSetupFailed(
            "libawig.so is not built at %s; run __graft_entry__.build() (hipcc, gfx950)" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, argtypes in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = ctypes.c_int
    lib.aw_last_error.argtypes = []
    lib.aw_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def check(status):
    """Raise the exception class mapped from a nonzero status (cufftCheckStatus analogue,
    cufft.py:96-104)."""
    if status != AW_OK:
        msg = load().aw_last_error()
        msg = msg.decode(errors="replace") if msg else ""
        raise _EXCEPTIONS.get(status, awigError)(msg)


def call(name, *args):
    check(getattr(load(), name)(*args))


def exported_symbols():
    return list(_SIGNATURES) + ["aw_last_error"]
