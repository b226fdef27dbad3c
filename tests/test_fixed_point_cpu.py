"""NumPy model of the line fixed point that c64 step sequences store between passes
(int32 pairs m, x = m 2^e, e = (biased exponent of the line's largest |re|, |im|) - 1052),
and what it buys over complex64 storage: the codec's error bound, and the error growth of a
long pass sequence in both storages, which the 3e-7 bound of the GPU test separates."""
import numpy as np


def encode(lines):
    """lines [L, N] complex128 -> (m [L, N, 2] int32, e [L]) as the pass stores them."""
    mx = np.maximum(np.abs(lines.real), np.abs(lines.imag)).max(axis=1)
    _, E = np.frexp(mx)                      # mx in [2^(E-1), 2^E): biased exponent E - 1 + 1023
    e = np.where(mx > 0, E - 30, -1052)
    m = np.stack([np.rint(np.ldexp(lines.real, -e[:, None])), np.rint(np.ldexp(lines.imag, -e[:, None]))], axis=-1)
    assert np.abs(m).max() <= 2 ** 30
    return m.astype(np.int32), e


def decode(m, e):
    return np.ldexp(m[..., 0].astype(np.float64), e[:, None]) + 1j * np.ldexp(m[..., 1].astype(np.float64), e[:, None])


def test_codec_error_bound():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((64, 256)) + 1j * rng.standard_normal((64, 256))
    x *= 10.0 ** rng.uniform(-30, 30, (64, 1))   # lines of distant magnitudes
    x[3] = 0
    x[5, 17] = 1e3 * np.abs(x[5]).max()          # one dominant element
    m, e = encode(x)
    y = decode(m, e)
    mx = np.maximum(np.abs(x.real), np.abs(x.imag)).max(axis=1)
    err = np.maximum(np.abs((y - x).real), np.abs((y - x).imag)).max(axis=1)
    assert np.all(err <= 2.0 ** -30 * mx)
    assert np.all(y[3] == 0)
    # exact values round-trip exactly
    z = np.ldexp(rng.integers(-2 ** 29, 2 ** 29, (4, 64)).astype(np.float64), -40) * (1 + 1j)
    mz, ez = encode(z)
    assert np.array_equal(decode(mz, ez), z)


def _passes(W, n, store, seed=1):
    """2n + 1 passes of [FFT along one axis, unit phases, inverse FFT] alternating axes (the
    shape of a merged step sequence), each followed by ``store``."""
    rng = np.random.default_rng(seed)
    for p in range(2 * n + 1):
        ax = 1 if p % 2 else 0
        ph = np.exp(2j * np.pi * rng.random(W.shape[ax]))
        F = np.fft.fft(W, axis=ax, norm="ortho")
        F *= ph[:, None] if ax == 0 else ph[None, :]
        W = store(np.fft.ifft(F, axis=ax, norm="ortho"), ax)
    return W


def _fixed(W, ax):
    lines = W.T if ax == 0 else W                 # the pass's lines
    m, e = encode(lines)
    out = decode(m, e)
    return out.T if ax == 0 else out


def test_fixed_point_error_growth_below_complex64():
    """4001 passes on a 128^2 random state: per-pass complex64 storage ends near 1.5e-6,
    line fixed point below 1e-7; the GPU test's 3e-7 bound lies between them."""
    rng = np.random.default_rng(2)
    W0 = rng.standard_normal((128, 128)) + 1j * rng.standard_normal((128, 128))
    n = 2000
    exact = _passes(W0, n, lambda W, ax: W)
    c64 = _passes(W0, n, lambda W, ax: W.astype(np.complex64).astype(np.complex128))
    fx = _passes(W0, n, _fixed)
    rel = lambda a: np.linalg.norm(a - exact) / np.linalg.norm(exact)
    assert rel(c64) > 3e-7, rel(c64)
    assert rel(fx) < 1e-7, rel(fx)
