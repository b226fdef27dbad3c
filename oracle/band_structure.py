"""Exact diagonalisation in mutually unbiased bases (TEST ORACLE).

Test infrastructure only (imported by ``tests/``).  Restates MUBQBandStructure
(band_structure_1d.py:7-92): the Hamiltonian H(k) = K(P + k) + V(X) of a periodic cell of
X_gridDIM points on [-X_amplitude, X_amplitude), written in the coordinate basis by sending
the diagonal momentum matrix through the sign-flipped DFTs (:64-80), and its lowest
eigenvalues (:82-92).  Used as the exact reference for Bloch <H>(beta), ground-state
energies and Gibbs states.

Pinned by tests/test_oracle_band.py: a free particle's spectrum is {K(P_j + k)} exactly; the
harmonic oscillator's lowest levels are n + 1/2 and its <H>(beta) is coth(beta/2)/2; H(k) is
Hermitian; dropping the Bloch vector k moves the free spectrum.
"""
import numpy as np


class MUBBandStructure:
    def __init__(self, X_gridDIM, X_amplitude, V, K):
        self.X_gridDIM = N = int(X_gridDIM)
        self.X_amplitude = Xa = float(X_amplitude)
        self.V, self.K = V, K
        self.dX = 2. * Xa / N
        self.X_range = np.linspace(-Xa, Xa - self.dX, N)               # :55-56
        r = np.arange(N)
        self.sign_flip = (-1.) ** (r[:, None] + r[None, :])              # :58-59
        self.P_range = (np.arange(N, dtype=np.float64) - N / 2) * (2. * np.pi / (N * self.dX))   # :61-63

    def get_hamiltonian(self, k=0.0):
        """:64-80: S . ifft_0 . fft_1 . S applied to diag K(P + k), plus diag V(X)."""
        h = np.diag(self.K(self.P_range + k)).astype(np.complex128)
        h *= self.sign_flip
        h = np.fft.ifft(h, axis=0)
        h = np.fft.fft(h, axis=1)
        h *= self.sign_flip
        return h + np.diag(self.V(self.X_range))

    def get_band_structure(self, ks, n):
        """:82-92: the n lowest eigenvalues for each Bloch vector, shape (n, len(ks))."""
        return np.array([np.linalg.eigvalsh(self.get_hamiltonian(k))[:n] for k in ks]).T

    def thermal_energy(self, beta, k=0.0):
        """<H> = sum E exp(-beta E) / sum exp(-beta E) of H(k)."""
        E = np.linalg.eigvalsh(self.get_hamiltonian(k))
        w = np.exp(-beta * (E - E.min()))
        return float(np.sum(E * w) / np.sum(w))
