"""Host side of the strong-field driver: the HHG spectrum formula
(strong_field_physics_cuda_1d.py:334-358) and the incremental .npz writer."""
import numpy as np


def test_hhg_spectrum_peaks_at_harmonic():
    from accuratewigner_amd.strong_field import hhg_spectrum
    omega, dt, N = 0.057, 0.05, 8192
    t = dt * np.arange(N)
    omegas, s = hhg_spectrum(np.cos(3 * omega * t), dt, omega)
    assert abs(abs(omegas[np.argmax(s)]) - 3.0) < 2 * (2 * np.pi / (N * dt)) / omega
    assert s.max() == 1.0
    # the same as the explicit window-FFT-shift recipe
    w = 0.42 - 0.5 * np.cos(2 * np.pi * np.arange(N) / (N - 1)) + 0.08 * np.cos(4 * np.pi * np.arange(N) / (N - 1))
    ref = np.abs(np.fft.fftshift(np.fft.fft(w * np.cos(3 * omega * t)))) ** 2
    assert np.allclose(s, ref / ref.max(), rtol=1e-12, atol=1e-15)


def test_npz_writer_layout(tmp_path):
    from accuratewigner_amd.strong_field import NpzWriter
    path = str(tmp_path / "x.npz")
    with NpzWriter(path) as w:
        w.add("settings/dt", 0.01)
        w.add("settings/V", "X * X")
        w.add("frames/0/wigner", np.arange(6.).reshape(2, 3))
        try:
            w.add("settings/E", lambda t: t)
            raise AssertionError("object values must be rejected")
        except TypeError:
            pass
    z = np.load(path, allow_pickle=False)
    assert sorted(z.files) == ["frames/0/wigner", "settings/V", "settings/dt"]
    assert np.array_equal(z["frames/0/wigner"], np.arange(6.).reshape(2, 3))
    assert float(z["settings/dt"]) == 0.01 and str(z["settings/V"]) == "X * X"


def test_npz_writer_background_thread(tmp_path):
    """The background writer stores the members in the order they were added, and reports
    an error of its thread at the next call."""
    from accuratewigner_amd.strong_field import NpzWriter
    path = str(tmp_path / "b.npz")
    frames = [np.random.default_rng(k).normal(size=(64, 128)) for k in range(12)]
    with NpzWriter(path, background=True) as w:
        for k, a in enumerate(frames):
            w.add("frames/%d/wigner" % k, a)
            w.add("frames/%d/t" % k, np.float64(0.5 * k))
    z = np.load(path, allow_pickle=False)
    assert len(z.files) == 24
    for k, a in enumerate(frames):
        assert np.array_equal(z["frames/%d/wigner" % k], a)
        assert float(z["frames/%d/t" % k]) == 0.5 * k
    w = NpzWriter(str(tmp_path / "c.npz"), background=True)
    w._zf.close()                  # a closed archive: the thread's write fails
    w.add("x", np.zeros(3))
    try:
        w.close()
        raise AssertionError("the thread's error must be raised")
    except ValueError:
        pass
