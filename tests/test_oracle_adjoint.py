"""Pins of the torch FP64 restatement used as the autograd reference of the adjoint: its
forward value equals the NumPy oracle's, and its autograd gradient equals central finite
differences of its own forward value (CPU, small grid)."""
import numpy as np
import pytest

from oracle.awig_oracle import WignerMoyalOracle

torch = pytest.importorskip("torch")


def _setup():
    Nx = Np = 64
    Xa = Pa = 6.
    dt, steps, npieces = 0.005, 60, 6
    n = np.arange(steps)
    pieces = np.minimum(((2 * n + 3) * npieces) // (2 * steps), npieces - 1)
    V0 = lambda X: 0.25 * X ** 4 - 0.5 * X ** 2     # noqa: E731
    K = lambda P: 0.5 * P * P                       # noqa: E731
    W0 = lambda X, P: np.exp(-2. * (X + 1.) ** 2 - 0.5 * P * P) / np.pi   # noqa: E731
    E = np.random.default_rng(3).uniform(-0.1, 0.1, npieces)
    return dict(Nx=Nx, Xa=Xa, Np=Np, Pa=Pa, dt=dt, steps=steps, pieces=pieces, V0=V0, K=K, W0=W0), E


def test_torch_restatement_forward_equals_numpy_oracle():
    from oracle.torch_wm import driven_expectation_x
    a, E = _setup()
    J = float(driven_expectation_x(torch.tensor(E), **a))
    pieces, dt = a["pieces"], a["dt"]
    V = lambda X, t: a["V0"](X) + X * E[pieces[int(round(t / dt - 1.5))]]   # noqa: E731
    o = WignerMoyalOracle(X_gridDIM=a["Nx"], X_amplitude=a["Xa"], P_gridDIM=a["Np"], P_amplitude=a["Pa"],
                          dt=dt, V=V, K=lambda P, t: a["K"](P))
    o.set_wignerfunction(a["W0"])
    o.propagate(a["steps"])
    Jo = o.get_average(lambda X, P, t: X)
    assert abs(J - Jo) < 1e-12 * max(1.0, abs(Jo))
    # the field moves <x>: a planted zero field gives a different value
    J0 = float(driven_expectation_x(torch.zeros(len(E), dtype=torch.float64), **a))
    assert abs(J0 - J) > 1e-6


def test_torch_restatement_autograd_equals_finite_differences():
    from oracle.torch_wm import driven_expectation_x
    a, E = _setup()
    Et = torch.tensor(E, requires_grad=True)
    J = driven_expectation_x(Et, **a)
    J.backward()
    g = Et.grad.numpy()
    h = 1e-4
    for k in (0, 3, 5):
        ep, em = E.copy(), E.copy()
        ep[k] += h
        em[k] -= h
        fd = (float(driven_expectation_x(torch.tensor(ep), **a))
              - float(driven_expectation_x(torch.tensor(em), **a))) / (2 * h)
        assert abs(fd - g[k]) < 1e-7 * max(1e-3, np.max(np.abs(g))), (k, fd, g[k])
