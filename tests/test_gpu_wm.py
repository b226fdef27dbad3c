"""GPU parity of the Wigner-Moyal path (libawig through the drop-in class) against the NumPy
oracle restatement of wigner_moyal_cuda_1d.py."""
import numpy as np
import pytest

from oracle.awig_oracle import WignerMoyalOracle, kconst

pytestmark = pytest.mark.gpu


def _pair(Nx=128, Np=128, Xa=10., Pa=10., dt=0.01, V="0.5 * X * X", Vf=lambda X, t: 0.5 * X * X,
          ehrenfest=False, precision="fp64", **kw):
    from accuratewigner_amd import WignerMoyalCUDA1D
    extra = {}
    oextra = {}
    if ehrenfest:
        extra = dict(diff_K="P", diff_V="X")
        oextra = dict(diff_K=lambda X, P, t: P, diff_V=lambda X, P, t: X)
    g = WignerMoyalCUDA1D(X_gridDIM=Nx, X_amplitude=Xa, P_gridDIM=Np, P_amplitude=Pa, dt=dt, t=0.,
                          K="0.5 * P * P", V=V, precision=precision, **extra, **kw)
    o = WignerMoyalOracle(X_gridDIM=Nx, X_amplitude=Xa, P_gridDIM=Np, P_amplitude=Pa, dt=dt,
                          V=Vf, K=lambda P, t: 0.5 * P * P, **oextra)
    return g, o


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("big", [False, True])
@pytest.mark.parametrize("n", [64, 128, 256, 512, 1024, 2048])
@pytest.mark.parametrize("prec", ["fp64", "mixed", "c64"])
def test_fft_axis_matches_numpy(n, prec, big):
    """Both kernel configurations: small batches (few lines per pass, latency configuration:
    up to 1024 lines, or 2048 of at most 256 points / 512 in FP32) and batches above that
    (throughput configuration)."""
    import torch
    from accuratewigner_amd.engine import Plan, FWD, INV
    from accuratewigner_amd.device import to_device
    rows, cols = (n, 128) if n != 128 else (128, 256)
    batch = 2048 // min(rows, cols) + 1 if big else 2
    rng = np.random.default_rng(n)
    a = rng.normal(size=(batch, rows, cols)) + 1j * rng.normal(size=(batch, rows, cols))
    dt = np.complex128 if prec == "fp64" else np.complex64
    # c64: FP64 arithmetic, so only the rounding of the stored input and output remains
    tol = {"fp64": 1e-13, "mixed": 2e-6, "c64": 3e-7}[prec]
    for axis, d in [(0, FWD), (0, INV), (1, FWD), (1, INV)]:
        p = Plan(rows, cols, batch, prec)
        x = to_device(a.astype(dt))
        p.fft(x, axis, d)
        got = x.get()
        ref = np.fft.fft(a, axis=axis + 1) if d == FWD else np.fft.ifft(a, axis=axis + 1, norm="forward")
        assert rel(got, ref) < tol, (axis, d, rel(got, ref))
    torch.cuda.synchronize()


def test_wigner_transforms_match_oracle():
    g, o = _pair(Nx=256, Np=128, Xa=8., Pa=6.)
    g.set_wignerfunction("exp(-pow(X - 1., 2) - pow(P - 0.5, 2)) / M_PI")
    o.set_wignerfunction(lambda X, P: np.exp(-(X - 1.) ** 2 - (P - 0.5) ** 2) / np.pi)
    assert rel(g.wignerfunction.get(), o.wignerfunction) < 1e-13
    assert rel(g.rho.get(), o.rho) < 1e-12
    g.rho2wigner_()
    o.rho2wigner_()
    assert rel(g.wignerfunction.get(), o.wignerfunction) < 1e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed", "c64"])
def test_wm_rho2wigner_blackman_matches_oracle(precision):
    """rho2wigner_blackman (wigner_moyal_cuda_1d.py:503-538) on a propagated state."""
    g, o = _pair(Nx=256, Np=128, Xa=8., Pa=6., precision=precision)
    g.set_wignerfunction("exp(-pow(X - 1., 2) - pow(P - 0.5, 2)) / M_PI")
    o.set_wignerfunction(lambda X, P: np.exp(-(X - 1.) ** 2 - (P - 0.5) ** 2) / np.pi)
    g.propagate(10)
    o.propagate(10)
    Wb = g.rho2wigner_blackman().get().astype(np.complex128)
    o.rho2wigner_(blackman=True)
    assert rel(Wb, o.wignerfunction) < {"fp64": 1e-12, "mixed": 2e-6, "c64": 5e-7}[precision]


def test_wm_propagate_fp64_matches_oracle():
    V = "0.5 * X * X + 0.1 * X * X * X * X"
    g, o = _pair(V=V, Vf=lambda X, t: 0.5 * X * X + 0.1 * X ** 4)
    g.set_wignerfunction("exp(-pow(X - 2., 2) - P * P) / M_PI")
    o.set_wignerfunction(lambda X, P: np.exp(-(X - 2.) ** 2 - P * P) / np.pi)
    W = g.propagate(300).get()
    Wo = o.propagate(300)
    assert rel(W, Wo) < 1e-10
    assert rel(g.rho.get(), o.rho) < 1e-10
    assert abs(g.t - o.t) == 0.0


def test_wm_config1_full_run_fp64():
    """Config 1: 256^2, x,p in [-10, 10), harmonic, W0 = exp(-(x-2)^2 - p^2)/pi, dt = 0.01,
    1000 steps, norm/<x>/<p>/<H> every step; FP64 rel-L2 vs the oracle <= 1e-10."""
    g, o = _pair(Nx=256, Np=256, ehrenfest=True)
    g.set_wignerfunction("exp(-pow(X - 2., 2) - P * P) / M_PI")
    o.set_wignerfunction(lambda X, P: np.exp(-(X - 2.) ** 2 - P * P) / np.pi)
    W = g.propagate(1000).get()
    Wo = o.propagate(1000)
    assert rel(W, Wo) < 1e-10
    for name in ("X_average", "P_average", "X_average_RHS", "P_average_RHS", "hamiltonian_average"):
        a, b = np.array(getattr(g, name)), np.array(getattr(o, name))
        assert a.shape == (1000,)
        assert np.max(np.abs(a - b)) < 1e-10, name
    t = 0.01 * np.arange(1, 1001)
    assert np.max(np.abs(np.array(g.X_average) - 2 * np.cos(t))) < 1e-3   # exact rotation


def test_wm_time_dependent_and_masks():
    V = "0.5 * X * X + X * 0.1 * sin(2. * t)"
    Vf = lambda X, t: 0.5 * X * X + X * 0.1 * np.sin(2. * t)  # noqa: E731
    ax = "pow(abs(sin(0.5 * M_PI * (XX + XX_max) / XX_max) * sin(0.5 * M_PI * (XX_prime + XX_prime_max) / XX_prime_max)), dt * 0.05)"
    g, o = _pair(V=V, Vf=Vf, abs_boundary_x=ax)
    XXm, XXpm = kconst(g.XX.max()), kconst(g.XX_prime.max())
    o.abs_boundary_x = lambda XX, XXp: np.abs(np.sin(0.5 * np.pi * (XX + XXm) / XXm)
                                              * np.sin(0.5 * np.pi * (XXp + XXpm) / XXpm)) ** (0.01 * 0.05)
    assert g._masks[0][0] == "vec"
    g.set_wignerfunction("exp(-pow(X - 2., 2) - P * P) / M_PI")
    o.set_wignerfunction(lambda X, P: np.exp(-(X - 2.) ** 2 - P * P) / np.pi)
    assert rel(g.propagate(150).get(), o.propagate(150)) < 1e-10


def _tmask_pair(ehrenfest):
    mx = "pow(fabs(cos(0.5 * M_PI * XX / 20.) * cos(0.5 * M_PI * XX_prime / 30.)), 0.02 * (1. + sin(t)))"
    mp = "exp(-0.001 * (1. + t) * (PP * PP + PP_prime * PP_prime))"
    fx = lambda XX, XXp, t: np.abs(np.cos(0.5 * np.pi * XX / 20.) * np.cos(0.5 * np.pi * XXp / 30.)) ** (0.02 * (1. + np.sin(t)))  # noqa: E731,E501
    fp = lambda PP, PPp, t: np.exp(-0.001 * (1. + t) * (PP * PP + PPp * PPp))  # noqa: E731
    V = "0.5 * X * X + 0.1 * X * sin(t)"
    Vf = lambda X, t: 0.5 * X * X + 0.1 * X * np.sin(t)  # noqa: E731
    g, o = _pair(Nx=128, Np=64, Xa=8., Pa=6., V=V, Vf=Vf, ehrenfest=ehrenfest, abs_boundary_x=mx, abs_boundary_p=mp)
    o.abs_boundary_x, o.abs_boundary_p = fx, fp
    g.set_wignerfunction("exp(-pow(X - 1., 2) - pow(P - 0.5, 2)) / M_PI")
    o.set_wignerfunction(lambda X, P: np.exp(-(X - 1.) ** 2 - (P - 0.5) ** 2) / np.pi)
    return g, o


@pytest.mark.parametrize("ehrenfest", [False, True])
def test_wm_time_dependent_separable_masks(ehrenfest):
    """Absorbing masks that depend on the step time (the reference evaluates the mask
    expression inside expV / expK with t in scope, wigner_moyal_cuda_1d.py:753, :777):
    per-step line and element vectors, merged steps and the Ehrenfest loop vs the oracle."""
    g, o = _tmask_pair(ehrenfest)
    assert g._tdep_masks()
    W = g.propagate(300 if not ehrenfest else 60).get()
    Wo = o.propagate(300 if not ehrenfest else 60)
    assert rel(W, Wo) < 1e-10
    if ehrenfest:
        for name in ("X_average", "P_average", "X_average_RHS", "P_average_RHS", "hamiltonian_average"):
            a, b = np.array(getattr(g, name)), np.array(getattr(o, name))
            assert np.max(np.abs(a - b)) < 1e-10 * max(1.0, np.max(np.abs(b))), name


def _ntmask_pair(ehrenfest, both, precision="fp64", batch=1):
    """A coordinate mask that depends on t and on XX - XX' (not separable) and a momentum
    mask that depends on t, separable (``both`` False) or not."""
    mx = "1. - 0.002 * (1. + sin(t)) * pow(XX - XX_prime, 2) / (1. + pow(XX - XX_prime, 2))"
    fx = lambda XX, XXp, t: 1. - 0.002 * (1. + np.sin(t)) * (XX - XXp) ** 2 / (1. + (XX - XXp) ** 2)  # noqa: E731
    if both:
        mp = "1. - 0.001 * t * pow(PP + PP_prime, 2) / (10. + pow(PP + PP_prime, 2))"
        fp = lambda PP, PPp, t: 1. - 0.001 * t * (PP + PPp) ** 2 / (10. + (PP + PPp) ** 2)  # noqa: E731
    else:
        mp = "exp(-0.001 * (1. + t) * (PP * PP + PP_prime * PP_prime))"
        fp = lambda PP, PPp, t: np.exp(-0.001 * (1. + t) * (PP * PP + PPp * PPp))  # noqa: E731
    V = "0.5 * X * X + 0.1 * X * sin(t)"
    Vf = lambda X, t: 0.5 * X * X + 0.1 * X * np.sin(t)  # noqa: E731
    kw = dict(batch=batch) if batch > 1 else {}
    g, o = _pair(Nx=128, Np=64, Xa=8., Pa=6., V=V, Vf=Vf, ehrenfest=ehrenfest, abs_boundary_x=mx,
                 abs_boundary_p=mp, precision=precision, **kw)
    o.abs_boundary_x, o.abs_boundary_p = fx, fp
    o.set_wignerfunction(lambda X, P: np.exp(-(X - 1.) ** 2 - (P - 0.5) ** 2) / np.pi)
    g.set_wignerfunction("exp(-pow(X - 1., 2) - pow(P - 0.5, 2)) / M_PI")
    return g, o


@pytest.mark.parametrize("ehrenfest,both", [(False, False), (False, True), (True, False), (True, True)])
def test_wm_time_dependent_nonseparable_masks(ehrenfest, both):
    """Masks that depend on t and are not separable are tabulated at every step's time, and
    the steps run unmerged: W and the Ehrenfest lists vs the oracle (FP64, 1e-10)."""
    g, o = _ntmask_pair(ehrenfest, both)
    assert g._masks[0][0] == "ttable" and g._masks[1][0] == ("ttable" if both else "time")
    steps = 300 if not ehrenfest else 260   # the Ehrenfest run crosses a full 256-step chunk
    W = g.propagate(steps).get()
    Wo = o.propagate(steps)
    assert rel(W, Wo) < 1e-10
    if ehrenfest:
        for name in ("X_average", "P_average", "X_average_RHS", "P_average_RHS", "hamiltonian_average"):
            a, b = np.array(getattr(g, name)), np.array(getattr(o, name))
            assert np.max(np.abs(a - b)) < 1e-10 * max(1.0, np.max(np.abs(b))), name
    # a later call continues from the same time; a single step evaluates the tables at t
    assert rel(g.propagate(7).get(), o.propagate(7)) < 1e-10
    g.single_step_propagation()
    o.single_step_propagation()
    assert rel(g.rho.get(), o.rho) < 1e-10


def test_wm_time_dependent_nonseparable_masks_mixed_and_batch():
    """The tabulated masks in mixed precision (FP32 tables) against FP64, and a batch of two
    identical states equal to one state, item by item."""
    g, o = _ntmask_pair(False, True, precision="mixed")
    assert rel(g.propagate(200).get(), o.propagate(200)) < 2e-5
    g2, o2 = _ntmask_pair(False, True, batch=2)
    W2 = g2.propagate(120).get()
    W1 = o2.propagate(120)
    for b in range(2):
        assert rel(W2[b], W1) < 1e-10


def test_wm_nonseparable_mask_table():
    ap = "(abs(PP + PP_prime) < 5. ? 1. : 0.98)"
    g, o = _pair(Nx=128, Np=64, abs_boundary_p=ap)
    o.abs_boundary_p = lambda PP, PPp: np.where(np.abs(PP + PPp) < 5., 1., 0.98)
    assert g._masks[1][0] == "table"
    g.set_wignerfunction("exp(-pow(X - 2., 2) - P * P) / M_PI")
    o.set_wignerfunction(lambda X, P: np.exp(-(X - 2.) ** 2 - P * P) / np.pi)
    assert rel(g.propagate(100).get(), o.propagate(100)) < 1e-10


def test_wm_observables_and_purity():
    g, o = _pair()
    g.set_wignerfunction("exp(-pow(X - 2., 2) - P * P) / M_PI")
    o.set_wignerfunction(lambda X, P: np.exp(-(X - 2.) ** 2 - P * P) / np.pi)
    g.propagate(50)
    o.propagate(50)
    assert abs(g.get_average("X * P + cos(X)") - o.get_average(lambda X, P, t: X * P + np.cos(X))) < 1e-12
    assert abs(g.get_purity() - o.get_purity()) < 1e-12
    assert abs(g.get_sigma_x_sigma_p() - o.get_sigma_x_sigma_p()) < 1e-11


def test_wm_mixed_precision_error_budget():
    g, o = _pair(precision="mixed")
    g.set_wignerfunction("exp(-pow(X - 2., 2) - P * P) / M_PI")
    o.set_wignerfunction(lambda X, P: np.exp(-(X - 2.) ** 2 - P * P) / np.pi)
    W = g.propagate(200).get()
    assert W.dtype == np.complex64
    err = rel(W.astype(np.complex128), o.propagate(200))
    assert err < 2e-5, err


def test_wm_c64_error_below_mixed():
    """complex64 states with FP64 arithmetic: only the stored state is rounded, in line fixed
    point (2^-30 of a line's largest value) inside the 200-step sequence and to complex64 at
    its end, so the error against the oracle stays near 1e-7 (4e-7 bound here; complex64 on
    every pass measured ~1e-6), far below the FP32-arithmetic tier's."""
    errs = {}
    for prec in ("c64", "mixed"):
        g, o = _pair(precision=prec, V="0.5 * X * X + 0.05 * X * X * X * X",
                     Vf=lambda X, t: 0.5 * X * X + 0.05 * X ** 4)
        g.set_wignerfunction("exp(-pow(X - 2., 2) - P * P) / M_PI")
        o.set_wignerfunction(lambda X, P: np.exp(-(X - 2.) ** 2 - P * P) / np.pi)
        W = g.propagate(200).get()
        assert W.dtype == np.complex64
        errs[prec] = rel(W.astype(np.complex128), o.propagate(200))
    assert errs["c64"] < 4e-7, errs
    assert errs["c64"] < 0.2 * errs["mixed"], errs


def test_wm_c64_throughput_configuration_matches_fp64():
    """1024^2 batches above the small-pass threshold (the throughput tiling of the FP64
    passes, states stored as complex64): 20 steps within 5e-7 of the FP64 path."""
    from accuratewigner_amd import WignerMoyalCUDA1D
    out = {}
    for prec in ("fp64", "c64"):
        s = WignerMoyalCUDA1D(X_gridDIM=1024, X_amplitude=10., P_gridDIM=1024, P_amplitude=10., dt=0.01, t=0.,
                              K="0.5 * P * P", V="0.5 * X * X + 0.05 * X * X * X * X", precision=prec, batch=3)
        X, P = s.X, s.P
        W0 = np.stack([np.exp(-(X - x0) ** 2 - (P - 0.3) ** 2) / np.pi for x0 in (1.0, -1.5, 0.2)])
        s.set_wignerfunction(W0)
        out[prec] = s.propagate(20).get().astype(np.complex128)
        del s
    for k in range(3):
        assert rel(out["c64"][k], out["fp64"][k]) < 5e-7, k


def test_wm_large_batch_fp64_matches_oracle():
    """A batch above the small-pass threshold (10 x 256 lines) runs the throughput kernel
    configuration; every state matches the oracle."""
    from accuratewigner_amd import WignerMoyalCUDA1D
    kw = dict(X_gridDIM=256, X_amplitude=8., P_gridDIM=256, P_amplitude=8., dt=0.01)
    g = WignerMoyalCUDA1D(t=0., K="0.5 * P * P", V="0.5 * X * X + 0.1 * X * X * X * X",
                          precision="fp64", batch=10, **kw)
    x0 = np.linspace(-1.5, 1.5, 10)
    X, P = g.X, g.P
    W0 = np.exp(-(X[None] - x0[:, None, None]) ** 2 - (P[None] - 0.3) ** 2) / np.pi
    g.set_wignerfunction(W0)
    Wg = g.propagate(30).get()
    for k in (0, 4, 9):
        o = WignerMoyalOracle(V=lambda X, t: 0.5 * X * X + 0.1 * X ** 4, K=lambda P, t: 0.5 * P * P, **kw)
        o.set_wignerfunction(W0[k])
        o.propagate(30)
        assert rel(Wg[k], o.wignerfunction) < 1e-10, k


@pytest.mark.parametrize("precision", ["fp64", "mixed", "c64"])
def test_wm_batch_equals_separate_runs(precision):
    """B states in one launch give the same result as B separate single-state runs."""
    from accuratewigner_amd import WignerMoyalCUDA1D
    kw = dict(X_gridDIM=128, X_amplitude=8., P_gridDIM=64, P_amplitude=6., dt=0.01, t=0.,
              K="0.5 * P * P", V="0.5 * X * X + 0.1 * X * X * X * X", precision=precision)
    x0 = np.array([-1.0, 0.5, 2.0])
    gb = WignerMoyalCUDA1D(batch=3, **kw)
    X, P = gb.X, gb.P
    W0 = np.exp(-(X[None] - x0[:, None, None]) ** 2 - P[None] ** 2) / np.pi
    gb.set_wignerfunction(W0)
    Wb = gb.propagate(40).get()
    assert Wb.shape == (3, 64, 128)
    for k in range(3):
        g1 = WignerMoyalCUDA1D(**kw)
        g1.set_wignerfunction(W0[k])
        W1 = g1.propagate(40).get()
        assert np.array_equal(Wb[k], W1), k


def test_batched_2048_grid_mixed_matches_fp64():
    """2048^2 batches run the throughput configurations of the step passes (FP32 columns of
    2048 points exchange in real and imaginary halves): a short mixed-precision propagation
    stays within FP32 rounding of the FP64 one."""
    from accuratewigner_amd import WignerMoyalCUDA1D
    out = {}
    for prec in ("fp64", "mixed"):
        s = WignerMoyalCUDA1D(X_gridDIM=2048, X_amplitude=10., P_gridDIM=2048, P_amplitude=10., dt=0.01, t=0.,
                              K="0.5 * P * P", V="0.5 * X * X + 0.05 * X * X * X * X", precision=prec, batch=2)
        X, P = s.X, s.P
        W0 = np.stack([np.exp(-(X - x0) ** 2 - (P - 0.3) ** 2) / np.pi for x0 in (1.0, -1.5)])
        s.set_wignerfunction(W0)
        out[prec] = s.propagate(20).get().astype(np.complex128)
        del s
    assert out["fp64"].shape == (2, 2048, 2048)
    assert rel(out["mixed"], out["fp64"]) < 2e-5


@pytest.mark.parametrize("ehrenfest", [False, True])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_wm_static_factors_equal_per_step_factors(ehrenfest, prec):
    """With V and K free of t the factor vectors are built once and reused by every chunk
    (constant-step sequences); the result equals the per-step factor path bit for bit over
    several chunks and calls, masks included."""
    def run(static):
        g, _ = _pair(Nx=128, Np=64, Xa=8., Pa=6., V="0.5 * X * X + 0.01 * X * X * X * X",
                     ehrenfest=ehrenfest, precision=prec,
                     abs_boundary_x="pow(fabs(cos(0.5 * M_PI * XX / 20.)), 0.01)")
        assert g._static
        g._static = static
        g.set_wignerfunction("exp(-pow(X - 1., 2) - pow(P - 0.5, 2)) / M_PI")
        g.propagate(300)
        W = g.propagate(7).get()
        lists = [np.array(getattr(g, n)) for n in ("X_average", "P_average", "hamiltonian_average")] \
            if ehrenfest else []
        return W, lists
    Ws, Ls = run(True)
    Wd, Ld = run(False)
    assert np.array_equal(Ws, Wd)
    for a, b in zip(Ls, Ld):
        assert np.array_equal(a, b)



@pytest.mark.parametrize("prec,batch", [("fp64", 1), ("mixed", 1), ("fp64", 2)])
def test_wm_time_separable_observables_equal_per_step_transform(prec, batch):
    """Ehrenfest lists with a time-dependent V (a driving field): observables that depend on t
    through a few functions of t (diff_V = V0'(X) + E(t), K + V = K + V0 + X E(t)) are read
    from merged steps against time-independent weights; they equal the reference's per-step
    Wigner transform (obs_by_adjoint = False), over two chunks and a second call."""
    from accuratewigner_amd import WignerMoyalCUDA1D
    kw = dict(X_gridDIM=128, X_amplitude=12., P_gridDIM=128, P_amplitude=6., dt=0.01, t=0.,
              K="0.5 * P * P", V="-1. / sqrt(X * X + 1.37) + X * E(t)", precision=prec, batch=batch,
              functions="__device__ double E(double t) { return 0.05 * sin(0.3 * t) * cos(0.05 * t); }",
              diff_K="P", diff_V="X / pow(X * X + 1.37, 1.5) + E(t)",
              abs_boundary_x="pow(fabs(sin(0.5 * M_PI * (XX + XX_max) / XX_max) * sin(0.5 * M_PI * (XX_prime + XX_prime_max) / XX_prime_max)), 0.0005)")
    out = {}
    for fast in (True, False):
        g = WignerMoyalCUDA1D(**kw)
        g.obs_by_adjoint = fast
        W0 = np.exp(-(g.X - 1.) ** 2 - (g.P - 0.3) ** 2) / np.pi
        W1 = np.exp(-(g.X + 0.5) ** 2 - (g.P + 0.2) ** 2) / np.pi
        g.set_wignerfunction(np.stack([W0, W1]) if batch == 2 else W0)
        g.propagate(300)
        g.propagate(20)
        assert (g._ehr["tsep"] is not None) == fast
        out[fast] = (g.wignerfunction.get().astype(np.complex128),
                     {n: np.array(getattr(g, n)) for n in ("X_average", "P_average", "X_average_RHS",
                                                            "P_average_RHS", "hamiltonian_average")})
    tol = 1e-10 if prec == "fp64" else 2e-5
    (W1, l1), (W0_, l0) = out[True], out[False]
    assert np.linalg.norm(W1 - W0_) / np.linalg.norm(W0_) < tol
    for name in l0:
        assert l1[name].shape == l0[name].shape
        assert np.max(np.abs(l1[name] - l0[name])) < tol * max(1.0, np.max(np.abs(l0[name]))), name


def test_wm_non_separable_time_observable_falls_back():
    """An observable that is not a combination of a few time-independent functions (here
    cos(X t)) keeps the per-step Wigner transform, with the same lists as the oracle."""
    from accuratewigner_amd import WignerMoyalCUDA1D
    g = WignerMoyalCUDA1D(X_gridDIM=64, X_amplitude=8., P_gridDIM=64, P_amplitude=8., dt=0.01, t=0.,
                          K="0.5 * P * P", V="0.5 * X * X", diff_K="P", diff_V="X + 0.01 * cos(X * t)")
    o = WignerMoyalOracle(X_gridDIM=64, X_amplitude=8., P_gridDIM=64, P_amplitude=8., dt=0.01,
                          V=lambda X, t: 0.5 * X * X, K=lambda P, t: 0.5 * P * P,
                          diff_K=lambda X, P, t: P, diff_V=lambda X, P, t: X + 0.01 * np.cos(X * t))
    g.set_wignerfunction("exp(-pow(X - 1., 2) - pow(P, 2)) / M_PI")
    o.set_wignerfunction(lambda X, P: np.exp(-(X - 1.) ** 2 - P ** 2) / np.pi)
    g.propagate(40)
    o.propagate(40)
    assert g._ehr["tsep"] is None
    for name in ("X_average", "P_average", "X_average_RHS", "P_average_RHS", "hamiltonian_average"):
        a, b = np.array(getattr(g, name)), np.array(getattr(o, name))
        assert np.max(np.abs(a - b)) < 1e-10 * max(1.0, np.max(np.abs(b))), name


@pytest.mark.parametrize("case", ["static_V", "tdep_mask"])
def test_wm_time_separable_observables_other_cases(case):
    """The time-separable path with a time-independent V (constant factor vectors, only an
    observable depends on t) and with a time-dependent separable mask (per-step mask lines in
    the closing vectors): equal to the per-step Wigner transform."""
    from accuratewigner_amd import WignerMoyalCUDA1D
    kw = dict(X_gridDIM=128, X_amplitude=10., P_gridDIM=128, P_amplitude=8., dt=0.01, t=0.,
              K="0.5 * P * P", diff_K="P")
    if case == "static_V":
        kw.update(V="0.5 * X * X + 0.01 * X * X * X * X", diff_V="X + 0.04 * X * X * X + 0.001 * sin(t)",
                  abs_boundary_x="pow(fabs(cos(0.5 * M_PI * XX / 20.)), 0.002)")
    else:
        kw.update(V="0.5 * X * X + 0.1 * X * sin(t)", diff_V="X + 0.1 * sin(t)",
                  abs_boundary_x="pow(fabs(cos(0.5 * M_PI * XX / 20.) * cos(0.5 * M_PI * XX_prime / 30.)), "
                                 "0.02 * (1. + sin(t)))")
    out = {}
    for fast in (True, False):
        g = WignerMoyalCUDA1D(**kw)
        g.obs_by_adjoint = fast
        assert g._static == (case == "static_V")
        g.set_wignerfunction("exp(-pow(X - 1., 2) - pow(P - 0.5, 2)) / M_PI")
        g.propagate(270)
        assert (g._ehr["tsep"] is not None) == fast
        out[fast] = (g.wignerfunction.get(), {n: np.array(getattr(g, n)) for n in (
            "X_average", "P_average", "X_average_RHS", "P_average_RHS", "hamiltonian_average")})
    (W1, l1), (W0, l0) = out[True], out[False]
    assert np.linalg.norm(W1 - W0) / np.linalg.norm(W0) < 1e-10
    for name in l0:
        assert np.max(np.abs(l1[name] - l0[name])) < 1e-10 * max(1.0, np.max(np.abs(l0[name]))), name
