"""Wigner-Moyal grid constants and the Wigner <-> rho transforms as fused passes, shared by
WignerMoyalCUDA1D and the serving slot pools.

Grids follow wigner_moyal_cuda_1d.py:110-143; the transforms are the reference's
wigner2rho (:422-469) and rho2wigner_ (:471-501), each as four fused passes.
"""
import numpy as np

from .engine import FWD, INV
from .rtc import kconst

SQRT2 = np.sqrt(2.0)


class WMGrid:
    """Step sizes and kernel coordinates of a Wigner-Moyal grid (wigner_moyal_cuda_1d.py:110-143)."""

    def __init__(self, Nx, Xa, Np, Pa):
        self.Nx, self.Np, self.Xa, self.Pa = int(Nx), int(Np), float(Xa), float(Pa)
        self.dX = 2. * Xa / Nx
        self.dP = 2. * Pa / Np
        self.dXdP = self.dX * self.dP
        X = np.linspace(-Xa, Xa - self.dX, Nx)
        P = np.linspace(-Pa, Pa - self.dP, Np)
        PP_prime = SQRT2 * P
        self.dPP_prime = PP_prime[1] - PP_prime[0]
        XX_prime = (np.arange(Np, dtype=np.float64) - Np // 2) * (np.pi / (SQRT2 * Pa))
        self.dXX_prime = XX_prime[1] - XX_prime[0]
        XX = SQRT2 * X
        self.dXX = XX[1] - XX[0]
        PP = (np.arange(Nx, dtype=np.float64) - Nx // 2) * (np.pi / (SQRT2 * Xa))
        self.dPP = PP[1] - PP[0]
        self.X, self.P, self.XX, self.XX_prime, self.PP, self.PP_prime = X, P, XX, XX_prime, PP, PP_prime
        j = np.arange(Nx, dtype=np.float64)
        i = np.arange(Np, dtype=np.float64)
        # coordinates as the reference kernels recompute them from their #define'd steps
        # (e.g. XX = dXX * (j - 0.5 * X_gridDIM), wigner_moyal_cuda_1d.py:769-770)
        self.k_dXX, self.k_dXX_prime = kconst(self.dXX), kconst(self.dXX_prime)
        self.k_dPP, self.k_dPP_prime = kconst(self.dPP), kconst(self.dPP_prime)
        self.kXX = self.k_dXX * (j - 0.5 * Nx)
        self.kXXp = self.k_dXX_prime * (i - 0.5 * Np)
        self.kPP = self.k_dPP * (j - 0.5 * Nx)
        self.kPPp = self.k_dPP_prime * (i - 0.5 * Np)
        self.kX = kconst(self.dX) * (j - 0.5 * Nx)
        self.kP = kconst(self.dP) * (i - 0.5 * Np)


def shear_x(plan, g, data, dir_a, dir_b, conj=False, **kw):
    """phase_shearX (wigner_moyal_cuda_1d.py:639-654): exp(-i tan(pi/8) PP_j XX'_i) / Nx on rows
    (conj: the complex conjugate factor)."""
    c = -np.tan(np.pi / 8.)
    plan.run(data, 1, dir_a, dir_b, chirp=-c if conj else c, lstep=g.k_dXX_prime, lcenter=0.5 * g.Np,
             estep=g.k_dPP, ecenter=0.5 * g.Nx, scale=1.0 / g.Nx, **kw)


def shear_y(plan, g, data, dir_a, dir_b, conj=False, **kw):
    """phase_shearY (wigner_moyal_cuda_1d.py:656-670): exp(+i sin(pi/4) PP'_i XX_j) / Np on columns
    (conj: the complex conjugate factor)."""
    c = np.sin(np.pi / 4.)
    plan.run(data, 0, dir_a, dir_b, chirp=-c if conj else c, lstep=g.k_dXX, lcenter=0.5 * g.Nx,
             estep=g.k_dPP_prime, ecenter=0.5 * g.Np, scale=1.0 / g.Np, **kw)


def wigner2rho(plan, g, W, rho):
    """wigner_moyal_cuda_1d.py:422-469: [S, F0] [I1, SX, F1] [I0, SY, F0] [I1, SX, F1, S]."""
    plan.run(W, 0, FWD, 0, sign_in=1, out=rho)
    shear_x(plan, g, rho, INV, FWD)
    shear_y(plan, g, rho, INV, FWD)
    shear_x(plan, g, rho, INV, FWD, sign_out=1)


def rho2wigner(plan, g, rho, W, dscale=None, scale_rho=False):
    """wigner_moyal_cuda_1d.py:471-501: [S, F1, SX, I1] [F0, SY, I0] [F1, SX, I1] [I0, S]
    (unnormalised, like the reference; ``dscale``: complex per-item factors applied in the
    first pass, which with ``scale_rho`` also writes rho times them back into rho)."""
    shear_x(plan, g, rho, FWD, INV, sign_in=1, out=W, dscale=dscale, in_scaled_out=rho if scale_rho else None)
    shear_y(plan, g, W, FWD, INV)
    shear_x(plan, g, W, FWD, INV)
    plan.run(W, 0, INV, 0, sign_out=1)


def rho2wigner_adjoint(plan, g, data):
    """Conjugate transpose of rho2wigner, in place: with F^H = I (unnormalised DFTs) and
    D^H = conj(D) for the sign and shear factors, the reversed sequence is
    [S, F0] [F1, SX*, I1] [F0, SY*, I0] [F1, SX*, I1, S].  Re<a, L b> = Re<L^H a, b>."""
    plan.run(data, 0, FWD, 0, sign_in=1)
    shear_x(plan, g, data, FWD, INV, conj=True)
    shear_y(plan, g, data, FWD, INV, conj=True)
    shear_x(plan, g, data, FWD, INV, conj=True, sign_out=1)
