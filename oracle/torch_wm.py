"""Torch FP64 restatement of a driven Wigner-Moyal propagation (TEST ORACLE, autograd).

Test infrastructure only (imported by ``tests/``).  The forward map is the oracle's
WignerMoyalOracle sequence (wigner_moyal_cuda_1d.py:543-582 steps, :471-501 rho2wigner_,
:608-622 normalisation, :329-347 get_average) written with torch.fft on complex128 CPU
tensors so that autograd differentiates it with respect to piecewise-constant field
amplitudes E_k in V(X, t) = V0(X) + X E(t).  Step n evaluates V at tau_n = t0 + (n + 1.5) dt
(t advanced before the step, :571; kernels at t + dt/2, :776) and uses E[pieces[n]].

Pinned by tests/test_gpu_adjoint.py: its forward value equals the NumPy oracle's
(WignerMoyalOracle with the same field) to 1e-12, and its autograd gradient equals central
finite differences of that forward value.
"""
import numpy as np

from .awig_oracle import WignerMoyalOracle


def driven_expectation_x(E, Nx, Xa, Np, Pa, dt, steps, pieces, V0, K, W0):
    """J = <X>(T) for field amplitudes E (a float64 torch tensor, may require grad).
    V0(X), K(P): NumPy callables; W0(X, P): initial Wigner function on the kernel grid."""
    import torch

    o = WignerMoyalOracle(X_gridDIM=Nx, X_amplitude=Xa, P_gridDIM=Np, P_amplitude=Pa, dt=dt,
                          V=lambda X, t: V0(X), K=lambda P, t: K(P))
    o.set_wignerfunction(W0)
    c128 = torch.complex128
    F = lambda a, ax: torch.fft.fft(a, dim=ax)                   # noqa: E731
    Inv = lambda a, ax: torch.fft.ifft(a, dim=ax, norm="forward")  # noqa: E731
    S = torch.tensor(o.S, dtype=c128)
    kXX = torch.tensor(o.kXX)
    kXXp = torch.tensor(o.kXXp)
    V0j = torch.tensor(V0(o.kXX))
    V0i = torch.tensor(V0(o.kXXp))
    EK = torch.tensor(np.exp(-1j * dt * (K(o.kPP)[None, :] - K(o.kPPp)[:, None])), dtype=c128)
    rho = torch.tensor(o.rho, dtype=c128)
    for n in range(steps):
        e = E[int(pieces[n])]
        Vj = V0j + e * kXX
        Vi = V0i + e * kXXp
        phase = -0.5 * dt * (Vj[None, :] - Vi[:, None])
        EV = S * torch.polar(torch.ones_like(phase), phase)
        r = EV * rho
        r = F(Inv(r, 0), 1)
        r = EK * r
        r = Inv(F(r, 0), 1)
        r = r / (Nx * Np)
        rho = EV * r
    SX = torch.tensor(o.shearX_factor(), dtype=c128)
    SY = torch.tensor(o.shearY_factor(), dtype=c128)
    w = S * rho
    w = Inv(SX * F(w, 1), 1)
    w = Inv(SY * F(w, 0), 0)
    w = Inv(SX * F(w, 1), 1)
    w = S * Inv(w, 0)
    kX = torch.tensor(o.kX)
    norm = w.real.sum() * o.dXdP
    return (w.real * kX[None, :]).sum() * o.dXdP / norm
