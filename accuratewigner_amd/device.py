"""Device arrays and streams.

``DeviceArray`` is a thin stand-in for the PyCUDA ``GPUArray`` attributes the reference's
callers use (``.get()``, ``[:] =``, ``.copy()``, ``.fill()``, ``.shape``, ``.gpudata``; e.g.
strong_field_physics_cuda_1d.py:206, 250 and wigner_moyal_cuda_1d.py:394, 410, 954).  HBM is
allocated through PyTorch-ROCm (its caching allocator and current HIP stream); all compute
goes through libawig.
"""
import numpy as np

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as _t
        _torch = _t
    return _torch


def require_gpu():
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("accuratewigner_amd needs a ROCm GPU (MI355X, gfx950); none is visible")
    return t


def current_stream():
    """Handle of torch's current HIP stream (passed to libawig as hipStream_t)."""
    return torch().cuda.current_stream().cuda_stream


_NP2TORCH = None


def _dtype_map():
    global _NP2TORCH
    if _NP2TORCH is None:
        t = torch()
        _NP2TORCH = {
            np.dtype(np.complex128): t.complex128, np.dtype(np.complex64): t.complex64,
            np.dtype(np.float64): t.float64, np.dtype(np.float32): t.float32,
            np.dtype(np.int32): t.int32, np.dtype(np.int64): t.int64,
        }
    return _NP2TORCH


class DeviceArray:
    """A device array (HBM) with the GPUArray methods the reference's callers rely on."""

    def __init__(self, shape, dtype=np.complex128, tensor=None):
        t = require_gpu()
        self.dtype = np.dtype(dtype)
        if tensor is None:
            tensor = t.zeros(tuple(shape), dtype=_dtype_map()[self.dtype], device="cuda")
        self.tensor = tensor
        self.shape = tuple(tensor.shape)

    # -- GPUArray-like interface ---------------------------------------------------------
    @property
    def gpudata(self):
        return self.tensor.data_ptr()

    @property
    def ptr(self):
        return self.tensor.data_ptr()

    @property
    def size(self):
        return int(np.prod(self.shape))

    @property
    def nbytes(self):
        return self.size * self.dtype.itemsize

    def get(self):
        """Copy to a host numpy array (a D2H sync, like GPUArray.get())."""
        return self.tensor.cpu().numpy().copy()

    def set(self, value):
        self[:] = value
        return self

    def fill(self, value):
        self.tensor.fill_(complex(value) if self.dtype.kind == "c" else value)
        return self

    def copy(self):
        return DeviceArray(self.shape, self.dtype, tensor=self.tensor.clone())

    def __setitem__(self, idx, value):
        t = torch()
        if isinstance(value, DeviceArray):
            src = value.tensor
        elif isinstance(value, np.ndarray):
            src = t.from_numpy(np.ascontiguousarray(value.astype(self.dtype))).to("cuda")
        elif t.is_tensor(value):
            src = value
        else:
            self.tensor[idx] = complex(value) if self.dtype.kind == "c" else value
            return
        self.tensor[idx] = src.to(self.tensor.dtype)

    def __getitem__(self, idx):
        return DeviceArray(None, self.dtype, tensor=self.tensor[idx].contiguous())

    @property
    def real(self):
        return self.tensor.real

    def __len__(self):
        return self.shape[0]

    def __repr__(self):
        return "DeviceArray(shape=%s, dtype=%s)" % (self.shape, self.dtype)


def to_device(arr, dtype=None):
    """numpy -> DeviceArray (contiguous copy)."""
    t = require_gpu()
    a = np.ascontiguousarray(arr if dtype is None else np.asarray(arr, dtype=dtype))
    return DeviceArray(a.shape, a.dtype, tensor=t.from_numpy(a).to("cuda"))


def empty(shape, dtype):
    t = require_gpu()
    return DeviceArray(shape, dtype, tensor=t.empty(tuple(shape), dtype=_dtype_map()[np.dtype(dtype)],
                                                    device="cuda"))
