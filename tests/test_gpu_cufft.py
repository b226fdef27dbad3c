"""The cufft.py drop-in's plan-handle interface (cufftPlanMany / cufftExecZ2Z / cufftDestroy /
PlanZ2Z, as the reference's classes call them, cufft.py:107-127, :224-345) against numpy, and
the FFT engine (aw_fft_axis) against hipFFT through torch.fft on the same device arrays."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_cufft_plan_many_reference_layouts():
    from accuratewigner_amd import cufft
    from accuratewigner_amd.device import empty, to_device
    rng = np.random.default_rng(11)
    shape = (256, 128)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    # Plan_Z2Z_2D_Axis0 as the reference builds it: n, inembed, onembed from np.array
    # (int64 on this platform, read as int32 like cuFFT does), stride = columns, idist 1
    n = np.array([shape[0]])
    emb = np.array([shape[0], shape[1]])
    h0 = cufft.cufftPlanMany(1, n.ctypes.data, emb.ctypes.data, shape[1], 1, emb.ctypes.data, shape[1], 1,
                             cufft.CUFFT_Z2Z, shape[1])
    x = to_device(a)
    cufft.cufftExecZ2Z(h0, x.gpudata, x.gpudata, cufft.CUFFT_FORWARD)
    assert rel(x.get(), np.fft.fft(a, axis=0)) < 1e-14
    # Plan_Z2Z_2D_Axis1: stride 1, idist = columns, batch = rows; out of place
    n1 = np.array([shape[1]])
    h1 = cufft.cufftPlanMany(1, n1.ctypes.data, np.array(shape).ctypes.data, 1, shape[1],
                             np.array(shape).ctypes.data, 1, shape[1], cufft.CUFFT_Z2Z, shape[0])
    y = empty(shape, np.complex128)
    cufft.cufftExecZ2Z(h1, x.gpudata, y.gpudata, cufft.CUFFT_INVERSE)
    assert rel(y.get(), np.fft.ifft(np.fft.fft(a, axis=0), axis=1, norm="forward")) < 1e-14
    cufft.cufftDestroy(h0)
    cufft.cufftDestroy(h1)
    with pytest.raises(cufft.cufftInvalidPlan):
        cufft.cufftExecZ2Z(h1, x.gpudata, x.gpudata, cufft.CUFFT_FORWARD)
    # PlanZ2Z: rank 1 over contiguous batches, and rank 2
    p1 = cufft.PlanZ2Z((512,), batch=6)
    b = rng.standard_normal((6, 512)) + 1j * rng.standard_normal((6, 512))
    z = to_device(b)
    cufft.cufftExecZ2Z(p1.handle, z.gpudata, z.gpudata, cufft.CUFFT_FORWARD)
    assert rel(z.get(), np.fft.fft(b, axis=1)) < 1e-14
    p2 = cufft.PlanZ2Z((128, 256), batch=2)
    c = rng.standard_normal((2, 128, 256)) + 1j * rng.standard_normal((2, 128, 256))
    w = to_device(c)
    o = empty(c.shape, np.complex128)
    cufft.cufftExecZ2Z(p2.handle, w.gpudata, o.gpudata, cufft.CUFFT_FORWARD)
    assert rel(o.get(), np.fft.fft2(c)) < 1e-14
    assert rel(w.get(), c) == 0.0                       # out of place left the input alone
    with pytest.raises(cufft.cufftInvalidSize):
        cufft.cufftPlanMany(1, np.array([100], np.int32).ctypes.data, None, 1, 0, None, 1, 0, cufft.CUFFT_Z2Z, 4)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("n", [64, 128, 256, 512, 1024, 2048, 4096])
def test_fft_axis_matches_hipfft(n, prec):
    """aw_fft_axis against torch.fft (hipFFT / rocFFT) on the same device arrays, both axes and
    directions, in both kernel configurations (3 and 40 items of [n, 128])."""
    import torch
    from accuratewigner_amd.device import DeviceArray
    from accuratewigner_amd.engine import FWD, INV, Plan
    tdt = torch.complex128 if prec == "fp64" else torch.complex64
    tol = 1e-13 if prec == "fp64" else 2e-6
    for batch in (3, 40):
        p = Plan(n, 128, batch, prec)
        g = torch.Generator(device="cuda").manual_seed(n + batch)
        a = torch.randn((batch, n, 128), dtype=tdt, device="cuda", generator=g)
        for axis in (0, 1):
            for d in (FWD, INV):
                x = DeviceArray(None, np.complex128 if prec == "fp64" else np.complex64, tensor=a.clone())
                p.fft(x, axis, d)
                if d == FWD:
                    ref = torch.fft.fft(a, dim=axis + 1)
                else:
                    ref = torch.fft.ifft(a, dim=axis + 1, norm="forward")
                err = (torch.linalg.vector_norm(x.tensor - ref) / torch.linalg.vector_norm(ref)).item()
                assert err < tol, (n, batch, axis, d, err)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("n", [64, 256, 1024, 4096])
def test_cufft_real_transforms(n, prec):
    """Plan2DAxis0 / Plan2DAxis1 with the real types (cufft.py:129-217): cu_fft_D2Z keeps the
    n/2 + 1 non-negative frequencies of each line (numpy rfft), cu_ifft_D2Z is the same forward
    transform (libcufft's D2Z takes no direction), and cu_fft_Z2D / cu_ifft_Z2D give the n
    reals of the unnormalised inverse (numpy irfft times n); R2C / C2R in FP32; the handle
    functions cufftExecD2Z / cufftExecZ2D directly."""
    from accuratewigner_amd import cufft
    from accuratewigner_amd.device import empty, to_device
    fwd_t, inv_t = (cufft.CUFFT_D2Z, cufft.CUFFT_Z2D) if prec == "fp64" else (cufft.CUFFT_R2C, cufft.CUFFT_C2R)
    rdt, cdt = (np.float64, np.complex128) if prec == "fp64" else (np.float32, np.complex64)
    tol = 1e-13 if prec == "fp64" else 2e-6
    rng = np.random.default_rng(n)
    other = 128          # (a strided column plan's column count is a power of two as well)
    h = n // 2 + 1
    for axis, shape in ((1, (other, n)), (0, (n, other))):
        a = rng.standard_normal(shape)
        plan_cls = cufft.Plan2DAxis1 if axis == 1 else cufft.Plan2DAxis0
        pf, pi = plan_cls(shape, fwd_t), plan_cls(shape, inv_t)
        x = to_device(a.astype(rdt))
        hs = (other, h) if axis == 1 else (h, other)
        y = empty(hs, cdt)
        ref = np.fft.rfft(a, axis=axis)
        cufft.cu_fft_D2Z(x, y, pf)
        assert rel(y.get(), ref) < tol, (axis, "fwd")
        y2 = empty(hs, cdt)
        cufft.cu_ifft_D2Z(x, y2, pf)
        assert np.array_equal(y.get(), y2.get())
        # inverse from a Hermitian half spectrum whose zero and Nyquist terms carry imaginary
        # parts, which the complex-to-real transform ignores
        X = ref + 0.3j * (np.arange(h) == 0).reshape((1, h) if axis == 1 else (h, 1))
        z = to_device(X.astype(cdt))
        r = empty(shape, rdt)
        cufft.cu_ifft_Z2D(z, r, pi)
        back = np.fft.irfft(ref, n=n, axis=axis) * n
        assert rel(r.get(), back) < tol, (axis, "inv")
        r2 = empty(shape, rdt)
        cufft.cu_fft_Z2D(z, r2, pi)
        assert np.array_equal(r.get(), r2.get())
        assert rel(r.get() / n, a) < tol                 # round trip
        with pytest.raises(cufft.cufftInvalidType):
            cufft.cu_fft_Z2D(x, y, pf)                   # a forward plan does not run the inverse
    # the handle API on contiguous batches, out of place only
    hp = cufft.cufftPlanMany(1, np.array([n], np.int32).ctypes.data, None, 1, 0, None, 1, 0, fwd_t, 5)
    b = rng.standard_normal((5, n))
    xb, yb = to_device(b.astype(rdt)), empty((5, h), cdt)
    (cufft.cufftExecD2Z if prec == "fp64" else cufft.cufftExecR2C)(hp, xb.gpudata, yb.gpudata)
    assert rel(yb.get(), np.fft.rfft(b, axis=1)) < tol
    with pytest.raises(cufft.cufftInvalidType):
        cufft.cufftExecZ2Z(hp, xb.gpudata, xb.gpudata, cufft.CUFFT_FORWARD)
    with pytest.raises(cufft.cufftInvalidValue):
        (cufft.cufftExecD2Z if prec == "fp64" else cufft.cufftExecR2C)(hp, xb.gpudata, xb.gpudata)
    cufft.cufftDestroy(hp)
