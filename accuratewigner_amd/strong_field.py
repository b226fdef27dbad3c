"""Strong-field ionisation driver (strong_field_physics_cuda_1d.py:24-374) on the MI355X path.

The reference script builds a WignerBlochCUDA1D system for a soft-core atom in a laser
field, finds its ground state (cached in ``ground_state_rho.pickle``), propagates it in
frames of 100 steps with Ehrenfest lists, stores every frame's Wigner function, time and
density-matrix diagonal in ``strong_field_physics.hdf5`` and finally plots the high-harmonic
spectrum |FFT(blackman * <P>)|^2.

Here the same flow runs without plotting.  h5py is not available, so results go to a
``.npz`` archive that keeps the HDF5 layout as key paths (``settings/<name>``,
``settings/X_wigner``, ``settings/P_wigner``, ``frames/<n>/{wigner,t,prob}``,
``ehrenfest/{X_average,P_average,X_average_RHS,P_average_RHS,hamiltonian_average}``,
:113-125, :247-257, :368-373), written incrementally so the frames never sit in host
memory together.  The ground-state cache is ``ground_state_rho.npy`` (no pickle).
"""
import os
import queue
import threading
import zipfile

import numpy as np


def sys_params(X_gridDIM=2 * 1024, P_gridDIM=512, X_amplitude=150., P_amplitude=20., dt=0.01):
    """The reference's parameter set (strong_field_physics_cuda_1d.py:24-85); grid sizes may
    be reduced for tests."""
    return dict(
        t=0.,
        dt=dt,
        X_gridDIM=X_gridDIM,
        X_amplitude=X_amplitude,
        P_gridDIM=P_gridDIM,
        P_amplitude=P_amplitude,
        _gamma=0.01,
        omega=0.05698,
        F=0.04,
        Ip=0.59,
        functions="""
    // The laser field (on for 8 periods of the carrier)
    __device__ double E(double t)
    {{
        return -F * sin(omega * t) * pow(sin(omega * t / 16.), 2);
    }}
    """,
        abs_boundary_x="pow("
                       "    abs(sin(0.5 * M_PI * (XX + XX_max) / XX_max)"
                       "    * sin(0.5 * M_PI * (XX_prime + XX_prime_max) / XX_prime_max))"
                       ", dt * 0.05)",
        E=lambda self, t: -self.F * np.sin(self.omega * t) * np.sin(self.omega * t / 16.) ** 2,
        K="0.5 * P * P",
        diff_K="P",
        V="-1. / sqrt(X * X + 1.37) + X * E(t)",
        diff_V="X / pow(X * X + 1.37, 1.5) + E(t)",
    )


def rv_sys_params(X_gridDIM=1024, X_amplitude=200., dt=0.02, F=0.05, omega=0.06, cycles=4):
    """Von Neumann rho(x, x') in a laser pulse (the RhoVNeumannCUDA1D form of the driver):
    soft-core V = -1/sqrt(X^2 + 1) + X E(t), E(t) = F sin^2(pi t / T) cos(omega t) over
    T = cycles 2 pi / omega, and an absorbing edge over the outer 10% of the box,
    b(X) = cos(pi/2 (|X| - 0.9 X_amplitude) / (0.1 X_amplitude))^(1/8) (1 inside)."""
    return dict(
        t=0.,
        dt=dt,
        X_gridDIM=X_gridDIM,
        X_amplitude=X_amplitude,
        F=F,
        omega=omega,
        T_pulse=cycles * 2. * np.pi / omega,
        functions="""
    __device__ double E(double t)
    {
        const double s = sin(M_PI * t / T_pulse);
        return F * s * s * cos(omega * t);
    }
    """,
        abs_boundary_x="(fabs(X) < 0.9 * X_amplitude ? 1. : "
                       "pow(cos(0.5 * M_PI * (fabs(X) - 0.9 * X_amplitude) / (0.1 * X_amplitude)), 0.125))",
        K="0.5 * P * P",
        diff_K="P",
        V="-1. / sqrt(X * X + 1.) + X * E(t)",
        diff_V="X / pow(X * X + 1., 1.5) + E(t)",
    )


def rv_ground_state_rho(params, nsteps=5000):
    """rho0 = |psi0><psi0| of the field-free soft-core atom: SchrodingerWignerCUDA1D imaginary
    time (set_ground_state, schrodinger_wigner_cuda_1d.py:484-503) on the same grid, then
    rho[i, j] = psi(X_j) conj(psi(X_i)) (bra on i, ket on j)."""
    from .schrodinger_wigner_cuda_1d import SchrodingerWignerCUDA1D
    sw = SchrodingerWignerCUDA1D(X_gridDIM=params["X_gridDIM"], X_amplitude=params["X_amplitude"],
                                 dt=params["dt"], t=0., V="-1. / sqrt(X * X + 1.)", K="0.5 * P * P")
    sw.set_ground_state(nsteps)
    psi = sw.wavefunction.get().reshape(-1)
    return np.outer(np.conj(psi), psi)


class NpzWriter:
    """Incremental ``.npz`` writer: every ``add`` appends one ``<key>.npy`` member.

    With ``background=True`` the members are written by one writer thread, in the order they
    were added, so that writing a frame overlaps the next frame's propagation (the array
    handed to ``add`` must not be modified afterwards); ``close`` waits for the queue and
    raises any error the thread met."""

    def __init__(self, path, background=False):
        self.path = path
        self._zf = zipfile.ZipFile(path, "w", compression=zipfile.ZIP_STORED, allowZip64=True)
        self._q = self._thread = self._err = None
        if background:
            self._q = queue.Queue(maxsize=8)
            self._thread = threading.Thread(target=self._loop, name="npz-writer", daemon=True)
            self._thread.start()

    def _write(self, key, arr):
        with self._zf.open(key + ".npy", "w", force_zip64=True) as f:
            # (np.ascontiguousarray would turn a 0-d scalar into shape (1,))
            np.lib.format.write_array(f, arr if arr.flags.c_contiguous else np.array(arr, order="C"),
                                      allow_pickle=False)

    def _loop(self):
        while True:
            item = self._q.get()
            if item is None:
                return
            if self._err is None:
                try:
                    self._write(*item)
                except BaseException as e:    # reported by the next add / close
                    self._err = e

    def _check(self):
        if self._err is not None:
            err, self._err = self._err, None
            raise err

    def add(self, key, value):
        arr = np.asanyarray(value)
        if arr.dtype == object:
            raise TypeError("%s: not an array value" % key)
        if self._q is None:
            self._write(key, arr)
        else:
            self._check()
            self._q.put((key, arr))

    def close(self):
        if self._thread is not None:
            self._q.put(None)
            self._thread.join()
            self._thread = None
        self._zf.close()
        self._check()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def hhg_spectrum(P_average, dt, omega):
    """High-harmonic spectrum as in strong_field_physics_cuda_1d.py:334-358:
    |fftshift(FFT(blackman(N) <P>))|^2 normalised to its maximum, against frequency / omega."""
    J = np.asarray(P_average, dtype=np.float64)
    N = J.shape[0]
    spectrum = np.abs(np.fft.fftshift(np.fft.fft(np.blackman(N) * J))) ** 2
    omegas = np.fft.fftshift(np.fft.fftfreq(N, dt / (2 * np.pi))) / omega
    return omegas, spectrum / spectrum.max()


class StrongFieldRun:
    """VisualizeDynamicsPhaseSpace without the figure (strong_field_physics_cuda_1d.py:88-269)."""

    def __init__(self, params=None, writer=None, ground_state_file="ground_state_rho.npy",
                 num_iteration=100, abs_tol_purity=1e-10, precision="fp64", ground_state_max_steps=None):
        from .wigner_bloch_cuda_1d import WignerBlochCUDA1D
        self.sys_params = dict(params if params is not None else sys_params())
        self.writer = writer
        self.quant_sys = WignerBlochCUDA1D(precision=precision, **self.sys_params)
        qs = self.quant_sys
        # final time: 8 periods of the carrier (:188-197)
        self.T_final = 8 * 2 * np.pi / qs.omega
        self.num_iteration = int(num_iteration)
        self.num_frames = int(np.ceil(self.T_final / qs.dt / self.num_iteration))
        self.current_frame_num = 0
        self.times = [qs.t]
        # ground state, cached (:202-214)
        loaded = False
        if ground_state_file and os.path.exists(ground_state_file):
            rho = np.load(ground_state_file, allow_pickle=False)
            if rho.shape == qs.rho.shape:
                qs.rho[:] = rho
                loaded = True
        if not loaded:
            qs.get_ground_state(abs_tol_purity=abs_tol_purity, max_steps=ground_state_max_steps)
            if ground_state_file:
                np.save(ground_state_file, qs.rho.get(), allow_pickle=False)
        self.ground_state_loaded = loaded
        self.wigner_initial = qs.wignerfunction.get().real
        if writer is not None:
            for key, val in self.sys_params.items():
                try:
                    writer.add("settings/" + key, val)
                except (TypeError, ValueError):
                    pass      # functions and the like (h5py raises TypeError, :115-120)
            writer.add("settings/X_wigner", np.broadcast_to(qs.X, qs.X.shape))
            writer.add("settings/P_wigner", np.broadcast_to(qs.P, qs.P.shape))

    def frame(self):
        """One frame (:236-269): propagate num_iteration steps, store W, t and diag(rho).
        Re W is taken on the device and copied into page-locked memory (half the bytes of
        the complex array, one DMA); a background writer then stores it while the next frame
        runs."""
        import torch
        qs = self.quant_sys
        qs.propagate(self.num_iteration)
        w_dev = qs.wignerfunction.tensor.real
        host = torch.empty(w_dev.shape, dtype=w_dev.dtype, pin_memory=True)
        host.copy_(w_dev)
        wigner = host.numpy()
        prob = qs.rho.tensor.diagonal().cpu().numpy()
        if self.writer is not None:
            key = "frames/%d/" % self.current_frame_num
            self.writer.add(key + "wigner", wigner)
            self.writer.add(key + "t", np.float64(qs.t))
            self.writer.add(key + "prob", prob)
        self.current_frame_num += 1
        self.times.append(qs.t)
        return wigner, prob

    def run(self, max_frames=881):
        for _ in range(min(int(max_frames), self.num_frames)):
            self.frame()
        return self

    def save_ehrenfest(self):
        """:368-373."""
        qs = self.quant_sys
        if self.writer is not None:
            for name in ("X_average", "P_average", "X_average_RHS", "P_average_RHS", "hamiltonian_average"):
                self.writer.add("ehrenfest/" + name, np.asarray(getattr(qs, name), dtype=np.float64))

    def spectrum(self):
        qs = self.quant_sys
        return hhg_spectrum(qs.P_average, qs.dt, qs.omega)


def main(path="strong_field_physics.npz", max_frames=881):
    with NpzWriter(path, background=True) as w:
        run = StrongFieldRun(writer=w)
        run.run(max_frames)
        run.save_ehrenfest()
    return run


if __name__ == "__main__":
    main()
