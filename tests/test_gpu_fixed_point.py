"""Step sequences of complex64 states with FP64 arithmetic (precision "c64") hold the state
between their passes in line fixed point: int32 pairs with one power-of-two scale per line
(2^-30 of the line's largest component), rounded to complex64 once, at the end of the
sequence.  Checked against the same sequences on complex128 states."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _seq(prec, W0, n, seed=3):
    """n merged steps with random unit-modulus factor vectors (the 1/N of each FFT pair folded
    into one vector per pass) on the states W0 [B, N, N]."""
    from accuratewigner_amd.device import to_device
    from accuratewigner_amd.engine import Plan
    B, N, _ = W0.shape
    plan = Plan(N, N, B, prec)
    rng = np.random.default_rng(seed)

    def unit(shape, s=1.0):
        return s * np.exp(2j * np.pi * rng.random(shape))

    tabs = [unit((n + 1, N)), unit((n + 1, N), 1.0 / N), unit((n, N), 1.0 / N), unit((n, N))]
    d = to_device(W0.astype(plan.state_dtype))
    plan.steps(d, n, *(to_device(t.astype(plan.factor_dtype)) for t in tabs))
    return d.get().astype(np.complex128)


def _states(B, N, seed=11):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((B, N, N)) + 1j * rng.standard_normal((B, N, N))).astype(np.complex64)


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("N", [128, 1024, 2048, 4096])
def test_fixed_point_error_grows_from_2_to_the_minus_30(N):
    """4001 passes: complex64 storage on every pass would give about sqrt(4001) x 2.4e-8 =
    1.5e-6; line fixed point keeps the error below 3e-7 (bound stated here), with only the
    closing complex64 rounding (2^-24) at full size.  Rows of 2048 and 4096 points span
    several waves (the line maximum through LDS instead of a wave shuffle)."""
    n = {128: 2000, 1024: 200, 2048: 40, 4096: 10}[N]
    W0 = _states(2, N)
    ref = _seq("fp64", W0, n)
    got = _seq("c64", W0, n)
    for k in range(2):
        assert _rel(got[k], ref[k]) < 3e-7, (k, _rel(got[k], ref[k]))


def test_fixed_point_zero_and_distant_magnitudes():
    """A zero state stays exactly zero; a state scaled by 1e-25 keeps the relative error of
    the unscaled one (the scale follows each line)."""
    W = _states(3, 256)
    W[0] = 0
    W[2] = W[1] * np.float32(1e-25)
    ref = _seq("fp64", W, 30)
    got = _seq("c64", W, 30)
    assert np.all(got[0] == 0)
    e1, e2 = _rel(got[1], ref[1]), _rel(got[2], ref[2])
    assert e1 < 1e-7 and e2 < 1e-7, (e1, e2)
    assert abs(e2 / e1 - 1) < 0.5, (e1, e2)


def test_fixed_point_lines_of_zeros_beside_others():
    """States that vanish on most lines (a band of rows and one column set): per-line scales
    of zero lines decode to zero and the others keep their precision."""
    W = _states(2, 256)
    W[:, :100] = 0
    W[:, 140:] = 0
    W[1, :, ::3] = 0
    ref = _seq("fp64", W, 10)
    got = _seq("c64", W, 10)
    for k in range(2):
        assert _rel(got[k], ref[k]) < 1e-7, k


def test_fixed_point_nan_reaches_every_element():
    """One NaN in the input leaves no finite element, as on complex128 states (an FFT spreads
    a NaN over its line, and a line stored as non-finite decodes to Inf)."""
    W = _states(1, 128)
    W[0, 5, 7] = np.nan
    ref = _seq("fp64", W, 3)
    got = _seq("c64", W, 3)
    assert not np.isfinite(ref).any()
    assert not np.isfinite(got).any()
