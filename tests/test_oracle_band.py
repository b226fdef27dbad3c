"""Pins of the exact-diagonalisation oracle (oracle/band_structure.py) and the Wigner plot
normalisations.  CPU only."""
import numpy as np
import pytest

from oracle.band_structure import MUBBandStructure


def test_free_particle_spectrum_is_the_kinetic_grid():
    K = lambda p: 0.5 * p * p          # noqa: E731
    for k in (0.0, 0.13):
        b = MUBBandStructure(64, 5., lambda x: 0. * x, K)
        H = b.get_hamiltonian(k)
        assert np.max(np.abs(H - H.conj().T)) < 1e-12
        E = np.linalg.eigvalsh(H)
        assert np.max(np.abs(E - np.sort(K(b.P_range + k)))) < 1e-10 * np.max(E)


def test_harmonic_levels():
    b = MUBBandStructure(256, 10., lambda x: 0.5 * x * x, lambda p: 0.5 * p * p)
    E = b.get_band_structure([0.0], 6)[:, 0]
    assert np.max(np.abs(E - (np.arange(6) + 0.5))) < 1e-8
    # thermal energy of the oscillator: (1/2) coth(beta/2)
    assert abs(b.thermal_energy(1.3) - 0.5 / np.tanh(0.65)) < 1e-8


def test_planted_dropped_bloch_vector_breaks_the_free_spectrum():
    """A relabelling of the momenta (sign flips, FFT order) is a phase similarity that
    commutes with V and cannot be seen in a spectrum; the Bloch vector can: dropping k from
    K(P + k) moves the free spectrum."""
    b = MUBBandStructure(64, 5., lambda x: 0. * x, lambda p: 0.5 * p * p)
    ref = np.sort(0.5 * (b.P_range + 0.13) ** 2)
    assert np.max(np.abs(np.linalg.eigvalsh(b.get_hamiltonian(0.13)) - ref)) < 1e-10
    assert np.max(np.abs(np.linalg.eigvalsh(b.get_hamiltonian(0.0)) - ref)) > 1e-2


def test_wigner_normalize_maps_zero_to_half():
    from accuratewigner_amd.wigner_normalize import WignerNormalize, WignerSymLogNorm
    n = WignerNormalize(vmin=-0.01, vmax=0.1)
    v = n(np.array([-0.01, -0.005, 0., 0.05, 0.1]))
    assert np.allclose(v, [0., 0.25, 0.5, 0.75, 1.])
    assert float(n(0.)) == 0.5
    c = WignerNormalize(vmin=-1., vmax=1., clip=True)
    assert np.allclose(c(np.array([-5., 5.])), [0., 1.])
    m = n(np.ma.array([0., 0.1], mask=[False, True]))
    assert m.mask.tolist() == [False, True] and m[0] == 0.5
    s = WignerSymLogNorm(linthresh=1e-3, vmin=-0.1, vmax=0.1)
    w = s(np.array([-0.1, -1e-3, 0., 1e-3, 0.1]))
    assert np.allclose(w, [0., 0.5 - 0.5 / 3, 0.5, 0.5 + 0.5 / 3, 1.])
    assert np.all(np.diff(s(np.linspace(-0.1, 0.1, 101))) >= 0)
    with pytest.raises(ValueError):
        WignerNormalize(vmin=1., vmax=0.)(np.array([0.5]))
