"""Host-side wrapper of libawig plans: fused passes, factor vectors and reductions.

A ``Plan`` owns one ``aw_plan`` (twiddles + reduction scratch) for a batch of
[rows, cols] grids in one precision.  Everything runs on torch's current HIP stream and
never synchronises with the host unless a caller reads a result.
"""
import ctypes

import numpy as np

from . import _lib
from .device import DeviceArray, current_stream, empty

# "fp64": complex128 states and FP64 arithmetic (the reference's); "mixed": complex64 states
# and FP32 FFT / factor arithmetic; "c64": complex64 states with FP64 twiddles, factors, masks
# and arithmetic (only the stored state is rounded to FP32)
PRECISIONS = {"fp64": _lib.AW_PREC_F64, "mixed": _lib.AW_PREC_F32, "fp32": _lib.AW_PREC_F32,
              "c64": _lib.AW_PREC_C64}
STATE_DTYPE = {_lib.AW_PREC_F64: np.complex128, _lib.AW_PREC_F32: np.complex64, _lib.AW_PREC_C64: np.complex64}
# factor vectors, complex tables and (as the real dtype) masks follow the arithmetic
FACTOR_DTYPE = {_lib.AW_PREC_F64: np.complex128, _lib.AW_PREC_F32: np.complex64, _lib.AW_PREC_C64: np.complex128}
REAL_DTYPE = {_lib.AW_PREC_F64: np.float64, _lib.AW_PREC_F32: np.float32, _lib.AW_PREC_C64: np.float64}

FWD, INV = _lib.AW_FORWARD, _lib.AW_INVERSE


def _p(x):
    """Device pointer of a DeviceArray / tensor / int / None."""
    if x is None:
        return None
    if isinstance(x, DeviceArray):
        return x.ptr
    if isinstance(x, int):
        return x
    return x.data_ptr()


_DEFERRED_PLANS = []


def _capturing():
    """True while the current stream is being captured into a graph."""
    import torch
    return torch.cuda.is_available() and torch.cuda.is_current_stream_capturing()


def destroy_deferred_plans():
    """Destroy the plans that were dropped while a graph capture was open."""
    while _DEFERRED_PLANS and not _capturing():
        _lib.load().aw_plan_destroy(_DEFERRED_PLANS.pop())


class Plan:
    """Batched [rows, cols] engine (replaces the cuFFT plans Plan_Z2Z_2D_Axis0/1,
    cufft.py:274-345, and the PyCUDA elementwise kernels around them)."""

    def __init__(self, rows, cols, batch=1, precision="fp64"):
        self.rows, self.cols, self.batch = int(rows), int(cols), int(batch)
        self.prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        self.state_dtype = STATE_DTYPE[self.prec]
        self.factor_dtype = FACTOR_DTYPE[self.prec]
        self.real_dtype = REAL_DTYPE[self.prec]
        self.fsize = np.dtype(self.factor_dtype).itemsize   # bytes per factor-vector element
        h = ctypes.c_void_p()
        _lib.call("aw_plan_create", ctypes.byref(h), self.rows, self.cols, self.batch, self.prec)
        self.handle = h
        self._stats = empty((self.batch, 8), np.float64)

    def __del__(self):
        try:
            h = getattr(self, "handle", None)
            self.handle = None
            if h:
                if _capturing():
                    _DEFERRED_PLANS.append(h)      # no hipFree inside a graph capture
                else:
                    _lib.load().aw_plan_destroy(h)
        except Exception:
            pass

    # -- passes ------------------------------------------------------------------------
    def run(self, data, axis, dir_a=0, dir_b=0, line_fac=None, elem_fac=None, mask=None,
            line_bstride=0, elem_bstride=0, mask_bstride=0, sign_in=0, sign_out=0,
            chirp=0.0, lstep=0.0, lcenter=0.0, estep=0.0, ecenter=0.0, lfft=0, efft=0,
            scale=1.0, dscale=None, out=None, out_mode=0, stream=None, ctab=None, ctab_bstride=0,
            chirp_out=0.0, obs_weights=None, obs_nw=0, obs_partials=None, obs_mode=0, obs_norm2=0,
            obs_norm2_weights=None, in_scaled_out=None, obs_line_fac=None, obs_elem_fac=None, obs_diag_nw=0):
        ps = _lib.aw_pass()
        ps.axis, ps.dir_a, ps.dir_b = axis, dir_a, dir_b
        ps.sign_in, ps.sign_out, ps.out_mode = sign_in, sign_out, out_mode
        ps.line_fac, ps.line_fac_bstride = _p(line_fac), line_bstride
        ps.elem_fac, ps.elem_fac_bstride = _p(elem_fac), elem_bstride
        ps.mask, ps.mask_bstride = _p(mask), mask_bstride
        ps.chirp, ps.lstep, ps.lcenter, ps.estep, ps.ecenter = chirp, lstep, lcenter, estep, ecenter
        ps.lfftorder, ps.efftorder = lfft, efft
        ps.scale = scale
        ps.dscale = _p(dscale)
        ps.out = _p(out)
        ps.ctab, ps.ctab_bstride, ps.chirp_out = _p(ctab), ctab_bstride, chirp_out
        ps.obs_weights, ps.obs_nw, ps.obs_partials = _p(obs_weights), obs_nw, _p(obs_partials)
        ps.obs_mode, ps.obs_norm2, ps.obs_norm2_weights = obs_mode, int(obs_norm2), _p(obs_norm2_weights)
        ps.in_scaled_out = _p(in_scaled_out)
        ps.obs_line_fac, ps.obs_elem_fac, ps.obs_diag_nw = _p(obs_line_fac), _p(obs_elem_fac), int(obs_diag_nw)
        _lib.call("aw_pass_run", self.handle, ctypes.c_void_p(_p(data)), ctypes.byref(ps),
                  ctypes.c_void_p(stream if stream is not None else current_stream()))

    def steps(self, data, nsteps, col_line, col_elem, row_line, row_elem, col_line_bstride=0,
              col_elem_bstride=0, row_line_bstride=0, row_elem_bstride=0, mask_open_close=None,
              mask_mid=None, mask_row=None, mask_bstride=0, dscale=None, stream=None,
              const_steps=False, batch=0):
        """n merged Strang steps (2n + 1 passes) in one native call (aw_step_sequence)."""
        st = _lib.aw_steps()
        st.nsteps = nsteps
        st.const_steps = int(bool(const_steps))
        st.batch = int(batch)
        st.col_line, st.col_elem = _p(col_line), _p(col_elem)
        st.row_line, st.row_elem = _p(row_line), _p(row_elem)
        st.col_line_bstride, st.col_elem_bstride = col_line_bstride, col_elem_bstride
        st.row_line_bstride, st.row_elem_bstride = row_line_bstride, row_elem_bstride
        st.mask_open_close, st.mask_mid, st.mask_row = _p(mask_open_close), _p(mask_mid), _p(mask_row)
        st.mask_bstride = mask_bstride
        st.dscale = _p(dscale)
        _lib.call("aw_step_sequence", self.handle, ctypes.c_void_p(_p(data)), ctypes.byref(st),
                  ctypes.c_void_p(stream if stream is not None else current_stream()))

    def step_group(self, nsteps, batch=0):
        """Items per group that ``steps`` uses for nsteps steps over ``batch`` items (0: all):
        batches over 192 MiB run in groups that stay in the MALL between passes."""
        g = ctypes.c_int()
        _lib.call("aw_step_group", self.handle, int(nsteps), int(batch), ctypes.byref(g))
        return g.value

    def fft(self, data, axis, direction, out=None):
        _lib.call("aw_fft_axis", self.handle, ctypes.c_void_p(_p(data)),
                  ctypes.c_void_p(_p(out)), axis, direction, ctypes.c_void_p(current_stream()))

    # -- factors -----------------------------------------------------------------------
    def factors(self, out, n, nvec, A=None, ca=0.0, B=None, cb=0.0, c0=0.0, mag=None,
                mag_pow=1, scale=1.0, alt_sign=False, mode=0, a_stride=None, b_stride=None,
                mag_stride=None, out_offset=0, out_vstride=1):
        """out[out_offset + v * out_vstride][k] = scale * (-1)^k? * mag^pow * E(c0 + ca A + cb B),
        E = exp(i.) (mode 0) or exp(.) (mode 1), in units of length-n vectors; A, B, mag are
        FP64 device pointers (ints) or None."""
        esz = self.fsize
        _lib.call("aw_make_factors", self.handle, ctypes.c_void_p(_p(out) + out_offset * n * esz),
                  out_vstride * n, n, nvec, mode,
                  ctypes.c_void_p(A), n if a_stride is None else a_stride, ca,
                  ctypes.c_void_p(B), n if b_stride is None else b_stride, cb,
                  c0, ctypes.c_void_p(mag), 0 if mag_stride is None else mag_stride, mag_pow,
                  scale, int(bool(alt_sign)), ctypes.c_void_p(current_stream()))

    def new_factors(self, n, nvec):
        return empty((nvec, n), self.factor_dtype)

    # -- reductions --------------------------------------------------------------------
    def stats(self, data, out=None):
        """[batch, 8] device FP64: Re/Im sum, sum |z|^2, Re/Im sum z^2, Re/Im trace, 0."""
        out = out if out is not None else self._stats
        _lib.call("aw_reduce_stats", self.handle, ctypes.c_void_p(_p(data)), ctypes.c_void_p(_p(out)),
                  ctypes.c_void_p(current_stream()))
        return out

    def pair_dot(self, a, b, u, v, out):
        """out[batch, 4] (device FP64) = Re, Im sum conj(a) b u_j and Re, Im sum conj(a) b v_i."""
        _lib.call("aw_pair_dot", self.handle, ctypes.c_void_p(_p(a)), ctypes.c_void_p(_p(b)),
                  ctypes.c_void_p(_p(u)), ctypes.c_void_p(_p(v)), ctypes.c_void_p(_p(out)),
                  ctypes.c_void_p(current_stream()))
        return out

    def weighted_sums(self, data, weights, nw, out, factor=0.0, dscale=None, norm2=False):
        """out[batch, nw (+1)] (device FP64) = Re sum conj(weights[q]) data per item (weights:
        complex128 [nw, rows, cols]), and with norm2 also sum |data|^2 last; with dscale, also
        dscale = 1 / (out[:, 0] factor).  The first call on a plan allocates its scratch
        (never under graph capture)."""
        _lib.call("aw_weighted_sums", self.handle, ctypes.c_void_p(_p(data)), ctypes.c_void_p(_p(weights)),
                  int(nw), int(bool(norm2)), ctypes.c_void_p(_p(out)), float(factor), ctypes.c_void_p(_p(dscale)),
                  ctypes.c_void_p(current_stream()))
        return out

    def row_obs_blocks(self):
        """Blocks per item of the weighted-sum partials of a row step pass (run with obs_*)."""
        nb = ctypes.c_int()
        _lib.call("aw_row_obs_blocks", self.handle, ctypes.byref(nb))
        return nb.value

    def obs_blocks(self, axis):
        """Blocks per item of the weighted-sum partials of a step pass along ``axis``, or 0
        where that pass stores none (blocks of fewer than 64 threads)."""
        nb = ctypes.c_int()
        rc = _lib.load().aw_pass_obs_blocks(self.handle, int(axis), ctypes.byref(nb))
        return nb.value if rc == 0 else 0

    def weighted_partials(self, data, weights, nw, partials=None, norm2=False):
        """Block partials [batch, nblk, nw (+1)] of weighted_sums into ``partials`` (a device
        pointer or array); returns nblk (partials None: only the block count)."""
        nb = ctypes.c_int()
        _lib.call("aw_weighted_partials", self.handle, ctypes.c_void_p(_p(data)), ctypes.c_void_p(_p(weights)),
                  int(nw), int(bool(norm2)), ctypes.c_void_p(_p(partials)), ctypes.byref(nb),
                  ctypes.c_void_p(current_stream()))
        return nb.value

    def cyclic_diag_obs(self, data, xw, nx, pw_hat, np_, out, out_bstride, factor=0.0, dscale=None):
        """out[b, :1 + nx + np_] = trace, coordinate sums over diag rho (FP64 weights xw
        [nx, n]) and momentum sums Re sum_d c_d pw_hat[q, d] / n from the signed cyclic
        diagonal sums (pw_hat: complex128 [np_, n], DFTs of the momentum weights); with
        dscale, also dscale = 1 / (trace factor)."""
        _lib.call("aw_cyclic_diag_obs", self.handle, ctypes.c_void_p(_p(data)), ctypes.c_void_p(_p(xw)), int(nx),
                  ctypes.c_void_p(_p(pw_hat)), int(np_), ctypes.c_void_p(_p(out)), int(out_bstride),
                  float(factor), ctypes.c_void_p(_p(dscale)), ctypes.c_void_p(current_stream()))
        return out

    def cyclic_diag_sums(self, data, out):
        """out[batch, n] complex128 (device): (-1)^d sum_i data[i, (i + d) mod n]."""
        _lib.call("aw_cyclic_diag_sums", self.handle, ctypes.c_void_p(_p(data)), ctypes.c_void_p(_p(out)),
                  ctypes.c_void_p(current_stream()))
        return out

    def scale_from_stats(self, stats_ptr, stride, comp, mode, factor, dscale, batch=None):
        _lib.call("aw_scale_from_stats", ctypes.c_void_p(stats_ptr), stride, comp, mode, factor,
                  ctypes.c_void_p(_p(dscale)), self.batch if batch is None else batch,
                  ctypes.c_void_p(current_stream()))

    def scale(self, data, dscale=None, scale=1.0, n=None):
        n = self.rows * self.cols if n is None else n
        _lib.call("aw_scale", self.prec, ctypes.c_void_p(_p(data)), n, self.batch,
                  ctypes.c_void_p(_p(dscale)), scale, ctypes.c_void_p(current_stream()))


def finish_sums(partials, nblk, ncomp, batch, out, factor=0.0, dscale=None):
    """out[b, c] = sum of the nblk block partials of component c (one block per item, fixed
    order); with dscale, also dscale[b] = 1 / (out[b, 0] factor)."""
    _lib.call("aw_finish_sums", ctypes.c_void_p(_p(partials)), int(nblk), int(ncomp), int(batch),
              ctypes.c_void_p(_p(out)), float(factor), ctypes.c_void_p(_p(dscale)), ctypes.c_void_p(current_stream()))


def sum_partials(partials, nblk, ncomp, batch, out):
    _lib.call("aw_sum_partials", ctypes.c_void_p(_p(partials)), nblk, ncomp, batch,
              ctypes.c_void_p(_p(out)), ctypes.c_void_p(current_stream()))


def reduce_vec(prec, data, n, batch, out):
    _lib.call("aw_reduce_vec", prec, ctypes.c_void_p(_p(data)), n, batch, ctypes.c_void_p(_p(out)),
              ctypes.c_void_p(current_stream()))


def convert(src, src_prec, dst, dst_prec, n):
    _lib.call("aw_convert", ctypes.c_void_p(_p(src)), src_prec, ctypes.c_void_p(_p(dst)), dst_prec,
              n, ctypes.c_void_p(current_stream()))


def mem_info():
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    _lib.call("aw_mem_info", ctypes.byref(f), ctypes.byref(t))
    return f.value, t.value


def rank1_split(M, rtol=1e-13):
    """If the real table M[i, j] equals v[i] * u[j] (to rtol of max|M|), return (v, u), else
    None.  Used to turn a separable absorbing-mask expression into two vectors."""
    M = np.asarray(M, dtype=np.float64)
    amax = np.max(np.abs(M))
    if amax == 0.0:
        return np.zeros(M.shape[0]), np.zeros(M.shape[1])
    i0, j0 = np.unravel_index(np.argmax(np.abs(M)), M.shape)
    u = M[i0, :].copy()
    v = M[:, j0] / M[i0, j0]
    if np.max(np.abs(np.outer(v, u) - M)) <= rtol * amax:
        return v, u
    return None
