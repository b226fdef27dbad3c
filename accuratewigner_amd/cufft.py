"""Drop-in for the reference's cuFFT wrapper (cufft.py) on libawig's FFT engine.

Same public names: CUFFT_FORWARD / CUFFT_INVERSE, the cufftError hierarchy, cufftCheckStatus,
Plan_Z2Z_2D_Axis0 / Plan_Z2Z_2D_Axis1 (cufft.py:274-345) and fft_Z2Z / ifft_Z2Z
(cufft.py:233-238): batched, unnormalised complex128 FFTs along one axis of a 2D device
array, forward = exp(-2 pi i jk/N).  Arrays are ``DeviceArray`` (or any object with
``.gpudata`` and ``.shape``); complex64 arrays use the FP32 engine.
"""
import ctypes

import numpy as np

from . import _lib
from .device import current_stream

CUFFT_FORWARD = -1
CUFFT_INVERSE = 1
CUFFT_C2C = 0x29
CUFFT_Z2Z = 0x69


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
