"""Coherent strong-field driver (coherent_strong_field_physics_cuda_1d.py) on the MI355X path.

The reference script prepares a soft-core atom on a 32768-point line (x in [-2000, 2000)),
finds its 14 lowest stationary states by imaginary time with Gram-Schmidt, starts from the
ground state, and propagates it through an 8-cycle laser pulse in frames of 300 steps with
Ehrenfest lists.  Each frame shows the coordinate density and the momentum density of the
wave function with the bound states projected out.  At the end it reports how well <H> was
preserved (:322-326).

Here the same flow runs without plotting.  SchrodingerWignerCUDA1D handles the 32768-point
line with its two-level transform.  The projected momentum density (:246-254) is the DFT of
(-1)^n psi_proj(n), computed on the host per frame like the reference (fftpack.fft ->
numpy.fft).  Results go to an ``.npz`` archive in the HDF5 layout of the reference
(``settings/<name>``; the per-frame keys ``frames/<n>/{t, x_density, p_density}`` hold what
the reference plots; ``ehrenfest/...``).
"""
import numpy as np

from .strong_field import NpzWriter, hhg_spectrum  # noqa: F401  (the HHG spectrum of <P>, :334-358)


def coulomb_sys_params(X_gridDIM=16 * 2048, X_amplitude=2000., dt=0.01):
    """coherent_strong_field_physics_cuda_1d.py:24-85 (grid size may be reduced for tests)."""
    return dict(
        t=0.,
        dt=dt,
        X_gridDIM=X_gridDIM,
        X_amplitude=X_amplitude,
        omega=0.05698,
        F=0.06,
        Ip=0.59,
        functions="""
    // The laser field (on for 8 periods of the carrier)
    __device__ double E(const double t)
    {{
        return -F * sin(omega * t) * pow(sin(omega * t / 16.), 2);
    }}
    """,
        abs_boundary_x="(1. - exp(-50. * abs(X - X_amplitude) / X_amplitude)) * "
                       "(1. - exp(-50. * abs(X + X_amplitude) / X_amplitude))",
        E=lambda self, t: -self.F * np.sin(self.omega * t) * np.sin(self.omega * t / 16.) ** 2,
        K="0.5 * P * P",
        diff_K="P",
        V="-1. / sqrt(X * X + 1.37) + X * E(t)",
        diff_V="X / pow(X * X + 1.37, 1.5) + E(t)",
    )


def short_sys_params(**kw):
    """The short-range variant (:88-97)."""
    p = coulomb_sys_params(**kw)
    p.update(V="-1. / sqrt(1.37) * exp(-0.2246 * X * X) + X * E(t)")
    return p


def projected_momentum_density(psi_proj):
    """|DFT((-1)^n psi)|^2 normalised to its maximum (:246-254)."""
    w = np.asarray(psi_proj).reshape(-1) * (-1.) ** np.arange(psi_proj.size)
    rho_p = np.abs(np.fft.fft(w)) ** 2
    return rho_p / rho_p.max()


class CoherentRun:
    """VisualizeDynamics without the figure (coherent_strong_field_physics_cuda_1d.py:103-281)."""

    def __init__(self, params=None, writer=None, nstates=14, nsteps=10000, num_iteration=300,
                 precision="fp64"):
        from .schrodinger_wigner_cuda_1d import SchrodingerWignerCUDA1D
        self.params = dict(params if params is not None else coulomb_sys_params())
        self.writer = writer
        sysm = self.coulomb_sys = SchrodingerWignerCUDA1D(precision=precision, **self.params)
        # stationary states; the energy of the highest, then start from the ground state (:176-199)
        sysm.get_stationary_states(nstates, nsteps=nsteps)
        sysm.wavefunction[:] = sysm.stationary_states[-1]
        self.top_energy = sysm.get_energy()
        sysm.wavefunction[:] = sysm.stationary_states[0]
        self.Ip = -sysm.get_energy()
        self.E_kinetic = (0.5 * sysm.P ** 2 - self.Ip) / sysm.omega
        self.T_final = 8 * 2 * np.pi / sysm.omega
        self.num_iteration = int(num_iteration)
        self.num_frames = int(np.ceil(self.T_final / sysm.dt / self.num_iteration))
        self.current_frame_num = 0
        self.times = [sysm.t]
        if writer is not None:
            for key, val in self.params.items():
                try:
                    writer.add("settings/" + key, val)
                except (TypeError, ValueError):
                    pass

    def frame(self):
        """One frame (:228-281): propagate, coordinate density, projected momentum density."""
        sysm = self.coulomb_sys
        sysm.propagate(self.num_iteration)
        x_density = np.abs(sysm.wavefunction.get().reshape(-1)) ** 2
        proj = sysm.projectout_stationary_states(sysm.wavefunction.copy()).get()
        p_density = projected_momentum_density(proj)
        if self.writer is not None:
            key = "frames/%d/" % self.current_frame_num
            self.writer.add(key + "t", np.float64(sysm.t))
            self.writer.add(key + "x_density", x_density)
            self.writer.add(key + "p_density", p_density)
        self.current_frame_num += 1
        self.times.append(sysm.t)
        return x_density, p_density

    def run(self, max_frames=None):
        n = self.num_frames if max_frames is None else min(int(max_frames), self.num_frames)
        for _ in range(n):
            self.frame()
        return self

    def energy_preservation_percent(self):
        """(1 - min <H> / max <H>) * 100 (:322-326)."""
        h = np.array(self.coulomb_sys.hamiltonian_average)
        return (1. - h.min() / h.max()) * 100

    def save_ehrenfest(self):
        sysm = self.coulomb_sys
        if self.writer is not None:
            for name in ("X_average", "P_average", "X_average_RHS", "P_average_RHS", "hamiltonian_average"):
                self.writer.add("ehrenfest/" + name, np.asarray(getattr(sysm, name), dtype=np.float64))


def main(path="coherent_strong_field_physics.npz", max_frames=None):
    with NpzWriter(path) as w:
        run = CoherentRun(writer=w)
        run.run(max_frames)
        run.save_ehrenfest()
    print("\nHamiltonian is preserved within the accuracy of %f percent" % run.energy_preservation_percent())
    return run


if __name__ == "__main__":
    main()
