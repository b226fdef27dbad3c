"""Adjoint gradient of a driven Wigner-Moyal propagation with respect to its driving field.

Dynamics: WignerMoyalCUDA1D (wigner_moyal_cuda_1d.py:543-582) under
    V(X, t) = V0(X) + X E(t),
with E piecewise constant over ``n_pieces`` equal intervals of the propagation window.
Objective: J(E) = <X>(T), the average of X over the final Wigner function after the
reference's normalisation (normalize_rho_wigner + get_average("X"), :329-347, :608-622).

Step n (0-based) evaluates V at tau_n = t0 + (n + 1.5) dt: the reference advances t before
the step (:571) and its expV / expK kernels evaluate at t + dt/2 (:776, :751).  Step n
belongs to piece k(n) = floor((n + 1.5) n_pieces / steps), capped at n_pieces - 1.

Gradient.  A step is U_n = EV_n M EV_n with EV_n = S exp(-i dt/2 [V(XX_j) - V(XX'_i)]) and
the kinetic middle M = I1 F0 EK F1 I0 / (Nx Np).  Without absorbing masks every factor has
modulus one and U_n is unitary; its adjoint U_n^H is the same pass sequence with conjugated
factors (dt -> -dt at the same tau_n), i.e. a backward split-operator step.  With
lambda_N = dJ/d rho_N (J(rho + d) = J + Re<lambda, d>, <a, b> = sum conj(a) b) and
lambda_n = U_n^H lambda_{n+1},
    dJ/dE(tau_n) = g_n + g_{n+1},   g_b = (dt/2) Im sum_ij conj(lambda_b) rho_b (XX_j - XX'_i),
because dEV/dE = -i dt/2 (XX_j - XX'_i) EV and |EV| = 1.  lambda_N is the adjoint Wigner
transform of (X_j - J) dXdP / n applied to rho_N (n = Re sum W dXdP).

Memory: rho is checkpointed every ``checkpoint`` steps during the forward sweep; the backward
sweep recomputes each segment's states from its checkpoint, then runs the conjugated steps
and the pair reductions through it.
"""
import numpy as np

from . import graphs, wm_ops
from .device import DeviceArray, empty, to_device
from .engine import FWD, INV


def step_pieces(steps, n_pieces):
    """Piece index of each step: floor((n + 1.5) n_pieces / steps), capped at n_pieces - 1."""
    n = np.arange(steps, dtype=np.int64)
    return np.minimum(((2 * n + 3) * n_pieces) // (2 * steps), n_pieces - 1)


class WignerMoyalDipoleControl:
    """<X>(T) of a Wigner-Moyal propagation driven by X E(t), and its adjoint gradient."""

    def __init__(self, X_gridDIM, X_amplitude, P_gridDIM, P_amplitude, dt, steps, n_pieces, V0, W0,
                 K="0.5 * P * P", t=0., precision="fp64", checkpoint=50, **kwargs):
        from .wigner_moyal_cuda_1d import WignerMoyalCUDA1D
        if kwargs.get("abs_boundary_x") or kwargs.get("abs_boundary_p"):
            raise NotImplementedError("the adjoint assumes unitary steps: absorbing masks are not supported")
        self.sys = WignerMoyalCUDA1D(X_gridDIM=X_gridDIM, X_amplitude=X_amplitude, P_gridDIM=P_gridDIM,
                                     P_amplitude=P_amplitude, dt=dt, t=t, V=V0, K=K, precision=precision,
                                     **kwargs)
        self.steps, self.n_pieces = int(steps), int(n_pieces)
        self.checkpoint = int(checkpoint)
        if self.steps < 1 or self.n_pieces < 1 or self.checkpoint < 1:
            raise ValueError("steps, n_pieces and checkpoint must be positive")
        self.dt, self.t0 = float(dt), float(t)
        self.W0 = W0
        self.piece = step_pieces(self.steps, self.n_pieces)
        s = self.sys
        Nx, Np = s.X_gridDIM, s.P_gridDIM
        # time-independent FP64 tables: V0 on XX and XX', K on PP and PP' (hipRTC, :763-778)
        ts = to_device(np.array([self.t0]))
        tabs = []
        for which, x, n in ((0, s._dev_kXX, Nx), (0, s._dev_kXXp, Np), (1, s._dev_kPP, Nx), (1, s._dev_kPPp, Np)):
            out = empty((1, n), np.float64)
            s._eval_1d(which, x, n, ts, 1, 0.0, out)
            tabs.append(out)
        self._V0j, self._V0i = tabs[0].get()[0], tabs[1].get()[0]
        self._Kj, self._Ki = tabs[2], tabs[3]
        self._nck = (self.steps + self.checkpoint - 1) // self.checkpoint
        shape = (Np, Nx)
        self._ckpt = DeviceArray((self._nck + 1,) + shape, s.state_dtype)
        self._seg = DeviceArray((self.checkpoint + 1,) + shape, s.state_dtype)
        self._lam = DeviceArray(shape, s.state_dtype)
        self._u, self._v = to_device(s._kXX), to_device(s._kXXp)
        self._dots = empty((self.steps + 1, 4), np.float64)
        self._esz = np.dtype(s.state_dtype).itemsize
        self._state = None
        # persistent inputs of one backward segment (fixed addresses, so that the segment's
        # launches are captured once and replayed): V tables, the forward (+dt) and adjoint
        # (-dt) unmerged factor vectors, and the pair-dot rows of the segment
        K, p = self.checkpoint, s.plan
        self._segV = (empty((K, Nx), np.float64), empty((K, Np), np.float64))
        self._segF = {sg: (p.new_factors(Nx, 2 * K), p.new_factors(Np, 2 * K), p.new_factors(Np, K),
                           p.new_factors(Nx, K)) for sg in (1, -1)}
        self._segdots = empty((K + 1, 4), np.float64)
        self._graphs = {}

    # -- helpers ---------------------------------------------------------------------------
    def _check_field(self, E):
        E = np.asarray(E, dtype=np.float64).reshape(-1)
        if E.shape[0] != self.n_pieces:
            raise ValueError("E must hold %d piece amplitudes" % self.n_pieces)
        return E

    def _tables(self, E, n0, n1):
        """FP64 V tables of steps n0..n1-1: V0(XX_j) + E XX_j and V0(XX'_i) + E XX'_i."""
        e = E[self.piece[n0:n1]]
        s = self.sys
        Vj = self._V0j[None, :] + e[:, None] * s._kXX[None, :]
        Vi = self._V0i[None, :] + e[:, None] * s._kXXp[None, :]
        return to_device(Vj), to_device(Vi)

    def _factors(self, E, n0, n1, sign, merged):
        Vj, Vi = self._tables(E, n0, n1)
        f = self.sys._factors_from_tables(n1 - n0, Vj.ptr, Vi.ptr, self._Kj.ptr, self._Ki.ptr,
                                          sign * self.dt, merged=merged, k_stride=0)
        return f, (Vj, Vi)

    def _segment_factors(self, Vj_all, Vi_all, n0, n1):
        """Fill the persistent V tables and factor vectors of steps n0..n1-1 (both signs),
        copying the tables on the device from those of all steps."""
        s, m = self.sys, n1 - n0
        self._segV[0].tensor[:m].copy_(Vj_all.tensor[n0:n1])
        self._segV[1].tensor[:m].copy_(Vi_all.tensor[n0:n1])
        for sg in (1, -1):
            s._factors_from_tables(m, self._segV[0].ptr, self._segV[1].ptr, self._Kj.ptr, self._Ki.ptr,
                                   sg * self.dt, merged=False, k_stride=0, dst=self._segF[sg])

    def _segment_body(self, m, last):
        """The launches of one backward segment of m steps from the persistent buffers:
        recompute rho_{n0+1..n0+m} from the checkpoint in slot 0, then sweep lambda back
        through the segment with the adjoint steps, with the pair dots of each step (and of
        the segment's end state for the last segment) in the segment's dot rows."""
        s, p = self.sys, self.sys.plan
        esz = p.fsize                     # factor vectors (the states use self._esz)
        Nx, Np = s.X_gridDIM, s.P_gridDIM
        colL, colE, rowL, rowE = self._segF[1]
        aL, aE, bL, bE = self._segF[-1]
        for k in range(m):
            # rho_{n0+k+1} from rho_{n0+k}: [EV, I0] out of place, [F1, EK, I1], [F0, EV]
            p.run(self._slot(self._seg, k), 0, 0, INV, line_fac=colL.ptr + 2 * k * Nx * esz,
                  elem_fac=colE.ptr + 2 * k * Np * esz, out=self._slot(self._seg, k + 1))
            p.run(self._slot(self._seg, k + 1), 1, FWD, INV, line_fac=rowL.ptr + k * Np * esz,
                  elem_fac=rowE.ptr + k * Nx * esz)
            p.run(self._slot(self._seg, k + 1), 0, FWD, 0, line_fac=colL.ptr + (2 * k + 1) * Nx * esz,
                  elem_fac=colE.ptr + (2 * k + 1) * Np * esz)
        if last:
            p.pair_dot(self._lam, self._slot(self._seg, m), self._u, self._v, self._segdots.ptr + m * 32)
        for k in reversed(range(m)):
            # lambda_{n0+k} = U^H lambda_{n0+k+1}: the step with conjugated factors
            p.steps(self._lam, 1, aL.ptr + 2 * k * Nx * esz, aE.ptr + 2 * k * Np * esz,
                    bL.ptr + k * Np * esz, bE.ptr + k * Nx * esz)
            p.pair_dot(self._lam, self._slot(self._seg, k), self._u, self._v, self._segdots.ptr + k * 32)

    def _slot(self, arr, k):
        return arr.ptr + k * arr.shape[-2] * arr.shape[-1] * self._esz

    # -- forward ---------------------------------------------------------------------------
    def forward(self, E):
        """J(E) = <X>(T); keeps the checkpoints for ``gradient``."""
        E = self._check_field(E)
        s, p, K = self.sys, self.sys.plan, self.checkpoint
        s.t = self.t0
        s.set_wignerfunction(self.W0)
        self._ckpt.tensor[0].copy_(s.rho.tensor)
        # the V tables of all steps in one upload: a host-to-device copy waits for the work
        # already queued, so one per chunk would make the host and the GPU take turns
        Vj_all, Vi_all = self._tables(E, 0, self.steps)
        Nx, Np = s.X_gridDIM, s.P_gridDIM
        for c in range(self._nck):
            n0, n1 = c * K, min(self.steps, (c + 1) * K)
            colL, colE, rowL, rowE = s._factors_from_tables(n1 - n0, Vj_all.ptr + n0 * Nx * 8, Vi_all.ptr + n0 * Np * 8,
                                                            self._Kj.ptr, self._Ki.ptr, self.dt, merged=True,
                                                            k_stride=0)
            p.steps(s.rho, n1 - n0, colL, colE, rowL, rowE)
            self._ckpt.tensor[c + 1].copy_(s.rho.tensor)
        s.t = self.t0 + self.steps * self.dt
        s.rho2wigner_()
        norm = float(p.stats(s.wignerfunction).get()[0, 0]) * s.dXdP
        J = s.get_average("X") / norm
        self._state = (E.copy(), J, norm, (Vj_all, Vi_all))
        return J

    # -- backward --------------------------------------------------------------------------
    def gradient(self, E=None):
        """(J, dJ/dE) by the adjoint sweep; runs ``forward(E)`` first when E is given or
        differs from the last forward call."""
        if E is not None:
            E = self._check_field(E)
            if self._state is None or not np.array_equal(self._state[0], E):
                self.forward(E)
        if self._state is None:
            raise RuntimeError("call forward(E) first")
        E, J, norm, (Vj_all, Vi_all) = self._state
        s, p, K = self.sys, self.sys.plan, self.checkpoint
        g = s._grid
        # lambda_N = L^H f with f_ij = (X_j - J) dXdP / norm (L = rho2wigner)
        f = np.broadcast_to(((g.kX - J) * s.dXdP / norm)[None, :], (s.P_gridDIM, s.X_gridDIM))
        self._lam[:] = np.ascontiguousarray(f)
        wm_ops.rho2wigner_adjoint(p, g, self._lam)
        for c in reversed(range(self._nck)):
            n0, n1 = c * K, min(self.steps, (c + 1) * K)
            m, last = n1 - n0, c == self._nck - 1
            self._seg.tensor[0].copy_(self._ckpt.tensor[c])
            self._segment_factors(Vj_all, Vi_all, n0, n1)
            graphs.run_chunk(self._graphs, (m, last), lambda m=m, last=last: self._segment_body(m, last), True)
            rows = m + 1 if last else m
            self._dots.tensor[n0:n0 + rows].copy_(self._segdots.tensor[:rows])
        d = self._dots.get()
        gb = 0.5 * self.dt * (d[:, 1] - d[:, 3])
        per_step = gb[:-1] + gb[1:]
        grad = np.bincount(self.piece, weights=per_step, minlength=self.n_pieces)
        return J, grad


class DipoleExpectation:
    """torch.autograd.Function: J = <X>(T) of a WignerMoyalDipoleControl for a field tensor E;
    backward runs the adjoint sweep.  Usage: DipoleExpectation.apply(E, control)."""
    _fn = None

    @classmethod
    def apply(cls, E, control):
        if cls._fn is None:
            cls._fn = _make_function()
        return cls._fn.apply(E, control)


def _make_function():
    import torch

    class _Fn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, E, control):
            e = E.detach().to("cpu", torch.float64).numpy()
            J = control.forward(e)
            ctx.control = control
            ctx.field = e.copy()
            return torch.tensor(J, dtype=torch.float64, device=E.device)

        @staticmethod
        def backward(ctx, gJ):
            _, grad = ctx.control.gradient(ctx.field)
            g = torch.as_tensor(grad, dtype=torch.float64, device=gJ.device)
            return gJ * g, None

    return _Fn
