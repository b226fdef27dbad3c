"""Drop-in for the reference's cuFFT wrapper (cufft.py) on libawig's FFT engine.

Same public names: the CUFFT_* constants (cufft.py:6-14), the cufftError hierarchy and
cufftCheckStatus (:31-104), cufftPlanMany / cufftDestroy / cufftExecZ2Z (:107-127, :224-232),
PlanZ2Z (:243-271), Plan_Z2Z_2D_Axis0 / Plan_Z2Z_2D_Axis1 (:274-345) and fft_Z2Z / ifft_Z2Z
(:233-238): batched, unnormalised complex FFTs, forward = exp(-2 pi i jk/N).  Arrays are
``DeviceArray`` (or any object with ``.gpudata`` and ``.shape``); CUFFT_C2C plans (complex64)
use the FP32 engine.  The real transforms: Plan2DAxis0 / Plan2DAxis1 (:160-217) with
CUFFT_D2Z / CUFFT_Z2D (CUFFT_R2C / CUFFT_C2R in FP32), cu_fft_D2Z / cu_ifft_D2Z and
cu_fft_Z2D / cu_ifft_Z2D (:129-157), and cufftExecD2Z / cufftExecZ2D / cufftExecR2C /
cufftExecC2R on plan handles.

cufftPlanMany accepts the layouts the reference builds: rank 1 over contiguous batches
(istride 1, idist n), rank 1 over the strided columns of a 2D array (idist 1, istride = the
number of columns = batch), and contiguous rank-2 batches (complex types only).  A real
transform keeps the n/2 + 1 non-negative frequencies of each line (cuFFT's layout): rows of
n/2 + 1 complex values for contiguous batches, the first n/2 + 1 rows of the [n, batch]
array for strided columns.  As in libcufft, D2Z / R2C always transform forward and Z2D / C2R
inverse: the direction the wrapper passes (cufft.py:134, :138, :151, :156) is ignored.  Lengths are powers of two from 64
to 4096 (the column count of a strided plan as well); other layouts raise cufftInvalidSize,
other types cufftInvalidType.  n, inembed and onembed may be host addresses of int32 arrays
(as the reference passes them) or sequences.
"""
import ctypes

import numpy as np

from . import _lib
from .device import current_stream

CUFFT_FORWARD = -1
CUFFT_INVERSE = 1

CUFFT_R2C = 0x2a  # real to complex (interleaved)
CUFFT_C2R = 0x2c  # complex (interleaved) to real
CUFFT_C2C = 0x29  # complex to complex (interleaved)
CUFFT_D2Z = 0x6a  # double to double-complex (interleaved)
CUFFT_Z2D = 0x6c  # double-complex (interleaved) to double
CUFFT_Z2Z = 0x69  # double-complex to double-complex (interleaved)


class cufftError(_lib.awigError):
    """CUFFT error"""


class cufftInvalidPlan(cufftError):
    """CUFFT was passed an invalid plan handle."""


class cufftAllocFailed(cufftError):
    """CUFFT failed to allocate GPU memory."""


class cufftInvalidType(cufftError):
    """The user requested an unsupported type."""


class cufftInvalidValue(cufftError):
    """The user specified a bad memory pointer."""


class cufftInternalError(cufftError):
    """Internal driver error."""


class cufftExecFailed(cufftError):
    """CUFFT failed to execute an FFT on the GPU."""


class cufftSetupFailed(cufftError):
    """The CUFFT library failed to initialize."""


class cufftInvalidSize(cufftError):
    """The user specified an unsupported FFT size."""


class cufftUnalignedData(cufftError):
    """Input or output does not satisfy texture alignment requirements."""


cufftExceptions = {
    0x1: cufftInvalidPlan, 0x2: cufftAllocFailed, 0x3: cufftInvalidType, 0x4: cufftInvalidValue,
    0x5: cufftInternalError, 0x6: cufftExecFailed, 0x7: cufftSetupFailed, 0x8: cufftInvalidSize,
    0x9: cufftUnalignedData,
}


def cufftCheckStatus(status):
    """Raise the exception class mapped from a nonzero status (cufft.py:96-104)."""
    if status != 0:
        msg = _lib.load().aw_last_error()
        raise cufftExceptions.get(status, cufftError)(msg.decode(errors="replace") if msg else "")


class _AxisPlan:
    def __init__(self, shape, axis, fft_type=CUFFT_Z2Z):
        if len(shape) != 2:
            raise ValueError('invalid transform dimension')
        self.shape = tuple(int(s) for s in shape)
        self.axis = axis
        self.batch = self.shape[1] if axis == 0 else self.shape[0]
        self.fft_type = fft_type
        prec = _lib.AW_PREC_F64 if fft_type == CUFFT_Z2Z else _lib.AW_PREC_F32
        h = ctypes.c_void_p()
        cufftCheckStatus(_lib.load().aw_plan_create(ctypes.byref(h), self.shape[0], self.shape[1], 1, prec))
        self.handle = h

    def __del__(self):
        # like the reference, never complain on destruction (cufft.py:303-309)
        try:
            _lib.load().aw_plan_destroy(self.handle)
        except Exception:
            pass


class Plan_Z2Z_2D_Axis0(_AxisPlan):
    """FFT over axis 0 of a 2D array (cufft.py:274-309): one transform per column."""

    def __init__(self, shape, fft_type=CUFFT_Z2Z):
        _AxisPlan.__init__(self, shape, 0, fft_type)


class Plan_Z2Z_2D_Axis1(_AxisPlan):
    """FFT over axis 1 of a 2D array (cufft.py:312-345): one transform per row."""

    def __init__(self, shape, fft_type=CUFFT_Z2Z):
        _AxisPlan.__init__(self, shape, 1, fft_type)


def _exec(x, y, plan, direction):
    if tuple(x.shape) != plan.shape or tuple(y.shape) != plan.shape:
        raise cufftInvalidValue("array shape %s does not match plan %s" % (x.shape, plan.shape))
    want = np.complex128 if plan.fft_type == CUFFT_Z2Z else np.complex64
    if np.dtype(x.dtype) != want or np.dtype(y.dtype) != want:
        raise cufftInvalidType("plan type needs %s arrays" % np.dtype(want))
    out = int(y.gpudata) if int(y.gpudata) != int(x.gpudata) else None
    cufftCheckStatus(_lib.load().aw_fft_axis(plan.handle, ctypes.c_void_p(int(x.gpudata)),
                                             ctypes.c_void_p(out), plan.axis, direction,
                                             ctypes.c_void_p(current_stream())))


def fft_Z2Z(x_input_gpu, y_output_gpu, plan):
    """Forward unnormalised FFT (cufft.py:233-234)."""
    _exec(x_input_gpu, y_output_gpu, plan, CUFFT_FORWARD)


def ifft_Z2Z(x_input_gpu, y_output_gpu, plan):
    """Inverse unnormalised FFT (cufft.py:237-238)."""
    _exec(x_input_gpu, y_output_gpu, plan, CUFFT_INVERSE)


# ---------------------------------------------------------------------------------------
# the plan-handle interface (cufftPlanMany / cufftExecZ2Z / cufftDestroy)
# ---------------------------------------------------------------------------------------
class _ManyPlan:
    """What a cufftPlanMany handle stands for: an aw_plan and the passes one execution runs."""

    def __init__(self, rows, cols, batch, prec, axes):
        h = ctypes.c_void_p()
        cufftCheckStatus(_lib.load().aw_plan_create(ctypes.byref(h), rows, cols, batch, prec))
        self.handle, self.axes = h, axes

    def __del__(self):
        try:
            _lib.load().aw_plan_destroy(self.handle)
        except Exception:
            pass


_PLANS = {}
_NEXT_HANDLE = [1]


def _ints(v, count):
    """count ints from a host address of int32 values (how the reference passes n, inembed and
    onembed, e.g. ``n.ctypes.data``) or from a sequence; None stays None."""
    if v is None:
        return None
    if isinstance(v, (int, np.integer)):
        return list((ctypes.c_int32 * count).from_address(int(v)))
    return [int(x) for x in list(v)[:count]]


_REAL_TYPES = {CUFFT_D2Z: 1, CUFFT_R2C: 1, CUFFT_Z2D: 0, CUFFT_C2R: 0}   # 1: forward (real input)


def cufftPlanMany(rank, n, inembed, istride, idist, onembed, ostride, odist, fft_type, batch):
    """Create a batched FFT plan (cufft.py:107-116); returns the handle as a ctypes.c_uint."""
    if fft_type not in (CUFFT_Z2Z, CUFFT_C2C) and fft_type not in _REAL_TYPES:
        raise cufftInvalidType("unknown transform type 0x%x" % int(fft_type))
    prec = _lib.AW_PREC_F64 if fft_type in (CUFFT_Z2Z, CUFFT_D2Z, CUFFT_Z2D) else _lib.AW_PREC_F32
    dims = _ints(n, rank)
    real = fft_type in _REAL_TYPES
    # (a real transform's input and output distances differ, n reals against n/2 + 1 complex
    # values, and cuFFT ignores them without embeddings, as the reference's plans pass them)
    if rank not in (1, 2) or batch < 1 or (istride, idist if not real else 0) != (ostride, odist if not real else 0):
        raise cufftInvalidSize("unsupported plan: rank %d, batch %d" % (rank, batch))
    if real and rank != 1:
        raise cufftInvalidSize("real transforms: rank-1 plans only")
    ie = _ints(inembed, rank)
    oe = _ints(onembed, rank)
    if rank == 2 and ((ie and ie[1:] != dims[1:]) or (oe and oe[1:] != dims[1:])):
        raise cufftInvalidSize("rank-2 plans must be contiguous")
    if rank == 1 and istride == 1 and (idist in (0, dims[0]) or real):
        plan = _ManyPlan(1, dims[0], batch, prec, (1,))                  # rows of contiguous batches
    elif rank == 1 and idist == 1 and istride == batch:
        plan = _ManyPlan(dims[0], batch, 1, prec, (0,))                  # the columns of [n, batch]
    elif rank == 2 and istride == 1 and idist in (0, dims[0] * dims[1]):
        plan = _ManyPlan(dims[0], dims[1], batch, prec, (1, 0))          # contiguous 2D transforms
    else:
        raise cufftInvalidSize("unsupported layout: istride %d, idist %d" % (istride, idist))
    plan.fft_type = fft_type
    h = _NEXT_HANDLE[0]
    _NEXT_HANDLE[0] += 1
    _PLANS[h] = plan
    return ctypes.c_uint(h)


def _plan_of(plan):
    h = plan.value if isinstance(plan, ctypes.c_uint) else int(plan)
    p = _PLANS.get(h)
    if p is None:
        raise cufftInvalidPlan("invalid plan handle %d" % h)
    return h, p


def cufftDestroy(plan):
    """Destroy an FFT plan (cufft.py:119-127)."""
    h, _ = _plan_of(plan)
    del _PLANS[h]


def cufftExecZ2Z(plan, idata, odata, direction):
    """Execute a complex-to-complex plan on device addresses (cufft.py:224-232): in place when
    idata == odata.  A rank-2 plan transforms along the rows, then along the columns."""
    _, p = _plan_of(plan)
    if p.fft_type not in (CUFFT_Z2Z, CUFFT_C2C):
        raise cufftInvalidType("a real-transform plan runs through cufftExecD2Z / cufftExecZ2D")
    if direction not in (CUFFT_FORWARD, CUFFT_INVERSE):
        raise cufftInvalidValue("direction must be CUFFT_FORWARD or CUFFT_INVERSE")
    lib = _lib.load()
    src, dst = int(idata), int(odata)
    for k, axis in enumerate(p.axes):
        a = src if k == 0 else dst
        out = dst if dst != a else None
        cufftCheckStatus(lib.aw_fft_axis(p.handle, ctypes.c_void_p(a), ctypes.c_void_p(out), axis, direction,
                                         ctypes.c_void_p(current_stream())))


class PlanZ2Z:
    """Batched complex128 plan of rank len(shape) (cufft.py:243-271)."""

    def __init__(self, shape, batch=1):
        self.shape = shape
        self.batch = batch
        self.fft_type = CUFFT_Z2Z
        if len(self.shape) > 0:
            n = np.asarray(shape, np.int32)
            self.handle = cufftPlanMany(len(shape), n.ctypes.data, None, 1, 0, None, 1, 0, self.fft_type,
                                        self.batch)
        else:
            raise ValueError('invalid transform size')

    def __del__(self):
        try:
            cufftDestroy(self.handle)
        except Exception:
            pass


# ---------------------------------------------------------------------------------------
# real transforms (cufft.py:129-217)
# ---------------------------------------------------------------------------------------
def _exec_real(plan, idata, odata, forward, types):
    _, p = _plan_of(plan)
    if p.fft_type not in types:
        raise cufftInvalidType("plan type 0x%x does not run this transform" % p.fft_type)
    if int(idata) == int(odata):
        raise cufftInvalidValue("real transforms run out of place")
    fn = _lib.load().aw_fft_r2c if forward else _lib.load().aw_fft_c2r
    cufftCheckStatus(fn(p.handle, ctypes.c_void_p(int(idata)), ctypes.c_void_p(int(odata)), p.axes[0],
                        ctypes.c_void_p(current_stream())))


def cufftExecD2Z(plan, idata, odata, direction=CUFFT_FORWARD):
    """Double-to-double-complex forward FFT on device addresses (the n/2 + 1 non-negative
    frequencies of each line); the direction is ignored, as libcufft takes none."""
    _exec_real(plan, idata, odata, True, (CUFFT_D2Z,))


def cufftExecZ2D(plan, idata, odata, direction=CUFFT_INVERSE):
    """Double-complex-to-double inverse FFT (the half spectrum in, n reals out)."""
    _exec_real(plan, idata, odata, False, (CUFFT_Z2D,))


def cufftExecR2C(plan, idata, odata, direction=CUFFT_FORWARD):
    """FP32 counterpart of cufftExecD2Z."""
    _exec_real(plan, idata, odata, True, (CUFFT_R2C,))


def cufftExecC2R(plan, idata, odata, direction=CUFFT_INVERSE):
    """FP32 counterpart of cufftExecZ2D."""
    _exec_real(plan, idata, odata, False, (CUFFT_C2R,))


def _exec_real_arrays(x, y, plan, forward):
    real_in = plan.fft_type in (CUFFT_D2Z, CUFFT_R2C)
    if forward != real_in:
        raise cufftInvalidType("plan type 0x%x does not run this transform" % plan.fft_type)
    f = cufftExecD2Z if plan.fft_type == CUFFT_D2Z else cufftExecR2C if plan.fft_type == CUFFT_R2C else \
        cufftExecZ2D if plan.fft_type == CUFFT_Z2D else cufftExecC2R
    f(plan.handle, int(x.gpudata), int(y.gpudata))


def cu_fft_D2Z(x_input_gpu, y_output_gpu, plan):
    """Real-to-complex FFT through a D2Z plan (cufft.py:133-134)."""
    _exec_real_arrays(x_input_gpu, y_output_gpu, plan, True)


def cu_ifft_D2Z(x_input_gpu, y_output_gpu, plan):
    """cufft.py:137-138: passes CUFFT_INVERSE to cufftExecD2Z, which takes no direction, so it
    is the same forward real-to-complex FFT."""
    _exec_real_arrays(x_input_gpu, y_output_gpu, plan, True)


def cu_fft_Z2D(x_input_gpu, y_output_gpu, plan):
    """cufft.py:150-152: the complex-to-real inverse FFT (cufftExecZ2D takes no direction)."""
    _exec_real_arrays(x_input_gpu, y_output_gpu, plan, False)


def cu_ifft_Z2D(x_input_gpu, y_output_gpu, plan):
    """Complex-to-real inverse FFT through a Z2D plan (cufft.py:155-157)."""
    _exec_real_arrays(x_input_gpu, y_output_gpu, plan, False)


class _TypedAxisPlan:
    def __init__(self, shape, cuFFT_type, axis):
        if len(shape) != 2:
            raise ValueError('invalid transform dimension')
        self.shape = tuple(int(s) for s in shape)
        self.fft_type = cuFFT_type
        if axis == 0:
            self.batch = self.shape[1]
            n = np.array([self.shape[0]], np.int32)
            stride = self.shape[1]
            embed = np.array([self.shape[0], stride], np.int32)
            self.handle = cufftPlanMany(1, n.ctypes.data, embed.ctypes.data, stride, 1, embed.ctypes.data, stride, 1,
                                        self.fft_type, self.batch)
        else:
            self.batch = self.shape[0]
            n = np.array([self.shape[1]], np.int32)
            self.handle = cufftPlanMany(1, n.ctypes.data, None, 1, 0, None, 1, 0, self.fft_type, self.batch)

    def __del__(self):
        try:
            cufftDestroy(self.handle)
        except Exception:
            pass


class Plan2DAxis0(_TypedAxisPlan):
    """Plan of any type over axis 0 of a 2D array (cufft.py:160-185): one transform per column."""

    def __init__(self, shape, cuFFT_type):
        _TypedAxisPlan.__init__(self, shape, cuFFT_type, 0)


class Plan2DAxis1(_TypedAxisPlan):
    """Plan of any type over axis 1 of a 2D array (cufft.py:188-212): one transform per row."""

    def __init__(self, shape, cuFFT_type):
        _TypedAxisPlan.__init__(self, shape, cuFFT_type, 1)
