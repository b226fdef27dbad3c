"""CPU-only checks of the C-ABI library: it loads, exports every symbol include/awig.h
declares, maps status codes to the right exception classes, and hipRTC compiles the
generated physics sources for gfx950 (no GPU needed for compilation)."""
import os
import re

import pytest

from accuratewigner_amd import _lib, rtc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "awig.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(aw_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libawig.so not built")
    lib = _lib.load()
    names = _header_functions()
    assert "aw_pass_run" in names and "aw_plan_create" in names
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) <= set(_lib.exported_symbols())


def test_every_declared_function_has_ctypes_argument_types():
    """A declared entry point without argtypes would pass pointers as C ints."""
    missing = [n for n in _header_functions() if n not in _lib._SIGNATURES and n != "aw_last_error"]
    assert not missing, missing


def test_status_codes_raise_mapped_exceptions():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libawig.so not built")
    import ctypes
    lib = _lib.load()
    h = ctypes.c_void_p()
    with pytest.raises(_lib.awigInvalidSize):
        _lib.check(lib.aw_plan_create(ctypes.byref(h), 100, 128, 1, 0))  # not a power of two
    with pytest.raises(_lib.awigInvalidType):
        _lib.check(lib.aw_plan_create(ctypes.byref(h), 128, 128, 1, 7))
    with pytest.raises(_lib.awigInvalidPlan):
        _lib.check(lib.aw_plan_destroy(None))
    assert lib.aw_version() >= 100


def test_rtc_compiles_reference_style_expressions():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libawig.so not built")
    consts = rtc.cuda_consts(dict(X_gridDIM=256, dt=0.01, F=0.04, omega=0.05698, flag=True))
    assert "#define X_gridDIM 256\n" in consts and "#define dt 1.000000000000000e-02\n" in consts
    # strong_field_physics_cuda_1d.py:52-58 style functions block with doubled braces
    functions = """
    __device__ double E(double t)
    {{
        return -F * sin(omega * t) * pow(sin(omega * t / 16.), 2);
    }}
    """
    src = (rtc.PRELUDE + consts + functions
           + "__device__ double K(double P, double t) { return (0.5 * P * P); }\n"
           + "__device__ double V(double X, double t) { return (-1. / sqrt(X * X + 1.37) + X * E(t)); }\n"
           + rtc.EVAL_1D
           + r'''extern "C" __global__ void aw_init(double2 *W) {
                 const double X = 0.1 * threadIdx.x, X_prime = -0.2;
                 aw_state_store<double2>::put(W + threadIdx.x,
                     cuda_complex(exp(-0.5 * pow(X - 1., 2) + cuda_complex(0., 0.5) * X) *
                                  exp(-0.5 * pow(X_prime - 1., 2) - cuda_complex(0., 0.5) * X_prime)));
               }''')
    ok, log = rtc.check_source(src)
    assert ok, log
    bad_ok, bad_log = rtc.check_source(src + "\n__device__ double broken( { }\n")
    assert not bad_ok and "error" in bad_log.lower()


def test_entry_points_reject_null_and_bad_arguments_before_touching_the_gpu():
    """Argument checks run before any HIP call, so they hold on a machine without a GPU."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libawig.so not built")
    import ctypes
    lib = _lib.load()
    vp = ctypes.c_void_p
    steps = _lib.aw_steps()
    steps.nsteps = 1
    ps = _lib.aw_pass()
    cases = [
        (_lib.awigInvalidPlan, lambda: lib.aw_step_sequence(None, vp(1), ctypes.byref(steps), None)),
        (_lib.awigInvalidPlan, lambda: lib.aw_pass_run(None, vp(1), ctypes.byref(ps), None)),
        (_lib.awigInvalidPlan, lambda: lib.aw_row_obs_blocks(None, ctypes.byref(ctypes.c_int()))),
        (_lib.awigInvalidPlan, lambda: lib.aw_pass_obs_blocks(None, 0, ctypes.byref(ctypes.c_int()))),
        (_lib.awigInvalidPlan, lambda: lib.aw_step_group(None, 4, 0, ctypes.byref(ctypes.c_int()))),
        (_lib.awigInvalidPlan, lambda: lib.aw_pair_dot(None, vp(1), vp(1), vp(1), vp(1), vp(1), None)),
        (_lib.awigInvalidPlan, lambda: lib.aw_cyclic_diag_sums(None, vp(1), vp(1), None)),
        (_lib.awigInvalidPlan, lambda: lib.aw_weighted_sums(None, vp(1), vp(1), 1, 0, vp(1), 1., None, None)),
        (_lib.awigInvalidPlan, lambda: lib.aw_weighted_partials(None, vp(1), vp(1), 1, 0, None, None, None)),
        (_lib.awigInvalidPlan, lambda: lib.aw_cyclic_diag_obs(None, vp(1), vp(1), 3, vp(1), 3, vp(1), 8, 1.,
                                                               None, None)),
        (_lib.awigInvalidPlan, lambda: lib.aw_reduce_stats(None, vp(1), vp(1), None)),
        (_lib.awigInvalidPlan, lambda: lib.aw_make_factors(None, vp(1), 0, 8, 1, 0, None, 8, 0., None, 8, 0.,
                                                            0., None, 0, 1, 1., 0, None)),
        (_lib.awigInvalidValue, lambda: lib.aw_sum_partials(None, 1, 1, 1, vp(1), None)),
        (_lib.awigInvalidValue, lambda: lib.aw_finish_sums(vp(1), 1, 9, 1, vp(1), 1., None, None)),
        (_lib.awigInvalidValue, lambda: lib.aw_scale_from_stats(None, 8, 0, 0, 1., vp(1), 1, None)),
        (_lib.awigInvalidValue, lambda: lib.aw_project_out(0, vp(1), None, 1, 128, None)),
        (_lib.awigInvalidValue, lambda: lib.aw_project_out(0, vp(1), vp(1), 5000, 128, None)),
        (_lib.awigInvalidValue, lambda: lib.aw_reduce_vec(0, None, 8, 1, vp(1), None)),
        (_lib.awigInvalidType, lambda: lib.aw_reduce_vec(9, vp(1), 8, 1, vp(1), None)),
        (_lib.awigInvalidSize, lambda: lib.aw_plan_create(ctypes.byref(vp()), 64, 8192, 1, 0)),
        (_lib.awigInvalidSize, lambda: lib.aw_plan_create(ctypes.byref(vp()), 32, 64, 1, 0)),
    ]
    for exc, call in cases:
        with pytest.raises(exc):
            _lib.check(call())
        assert _lib.load().aw_last_error()          # a message is left for the caller


def test_cufft_plan_many_argument_checks_without_a_gpu_call():
    """Unknown transform types and rank-2 real transforms are refused before any GPU call."""
    from accuratewigner_amd import cufft
    import numpy as np
    with pytest.raises(cufft.cufftInvalidType):
        cufft.cufftPlanMany(1, np.array([64], np.int32).ctypes.data, None, 1, 0, None, 1, 0, 0x99, 1)
    for t in (cufft.CUFFT_R2C, cufft.CUFFT_C2R, cufft.CUFFT_D2Z, cufft.CUFFT_Z2D):
        with pytest.raises(cufft.cufftInvalidSize):
            cufft.cufftPlanMany(2, np.array([64, 64], np.int32).ctypes.data, None, 1, 0, None, 1, 0, t, 1)


def test_wm_physics_source_with_time_dependent_masks_compiles():
    """The Wigner-Moyal physics module (V, K tables, mask tables and per-step mask lines) with
    mask expressions in XX, XX', PP, PP' and t compiles with hipRTC (no GPU needed)."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libawig.so not built")
    import types
    from accuratewigner_amd import rtc
    from accuratewigner_amd.wigner_moyal_cuda_1d import WignerMoyalCUDA1D
    ns = types.SimpleNamespace(
        cuda_consts=rtc.cuda_consts(dict(X_gridDIM=64, P_gridDIM=64), None, None),
        K="0.5 * P * P", V="0.5 * X * X + 0.1 * X * sin(t)",
        abs_boundary_x="pow(fabs(cos(0.5 * M_PI * XX / 20.) * cos(0.5 * M_PI * XX_prime / 30.)), 0.02 * (1. + sin(t)))",
        abs_boundary_p="exp(-0.001 * (1. + t) * (PP * PP + PP_prime * PP_prime))")
    ok, log = rtc.check_source(WignerMoyalCUDA1D._physics_source(ns))
    assert ok, log
