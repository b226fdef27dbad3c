"""The reduction entry points of the C ABI against NumPy on random data (FP64 and mixed):
signed cyclic diagonal sums, weighted sums (with sum |z|^2 and the normalisation scale),
their block partials finished by aw_finish_sums, the partials a row step pass stores as it
writes its output, the von Neumann Ehrenfest sums, and the adjoint's pair dots."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rand(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _tol(prec):
    return 1e-12 if prec == "fp64" else 1e-5


@pytest.mark.parametrize("prec", ["fp64", "mixed", "c64"])
def test_cyclic_diag_sums(prec):
    from accuratewigner_amd.device import empty, to_device
    from accuratewigner_amd.engine import Plan
    n, B = 128, 2
    p = Plan(n, n, B, prec)
    z = _rand(np.random.default_rng(1), (B, n, n)).astype(p.state_dtype)
    out = empty((B, n), np.complex128)
    p.cyclic_diag_sums(to_device(z), out)
    i = np.arange(n)
    ref = np.array([[(-1) ** d * np.sum(z[b, i, (i + d) % n]) for d in range(n)] for b in range(B)])
    assert np.max(np.abs(out.get() - ref)) < _tol(prec) * np.max(np.abs(ref))


@pytest.mark.parametrize("prec", ["fp64", "mixed", "c64"])
def test_weighted_sums_partials_and_finish(prec):
    from accuratewigner_amd.device import empty, to_device
    from accuratewigner_amd.engine import Plan, finish_sums
    rows, cols, B, nw = 64, 256, 3, 5
    rng = np.random.default_rng(2)
    p = Plan(rows, cols, B, prec)
    z = _rand(rng, (B, rows, cols)).astype(p.state_dtype)
    w = _rand(rng, (nw, rows, cols))
    zd, wd = to_device(z), to_device(w)
    out = empty((B, nw + 1), np.float64)
    ds = empty((B, 2), np.float64)
    p.weighted_sums(zd, wd, nw, out, factor=0.25, dscale=ds, norm2=True)
    zz = z.astype(np.complex128)
    ref = np.concatenate([np.einsum("qij,bij->bq", np.conj(w), zz).real,
                          np.sum(np.abs(zz) ** 2, axis=(1, 2))[:, None]], axis=1)
    got = out.get()
    assert np.max(np.abs(got - ref)) < _tol(prec) * np.max(np.abs(ref))
    assert np.allclose(ds.get()[:, 0], 1.0 / (got[:, 0] * 0.25), rtol=1e-14, atol=0)
    assert np.all(ds.get()[:, 1] == 0)
    # the same sums from the block partials and one finish launch: bit-identical
    nblk = p.weighted_partials(None, None, nw, norm2=True)
    part = empty((B, nblk, nw + 1), np.float64)
    assert p.weighted_partials(zd, wd, nw, part, norm2=True) == nblk
    out2 = empty((B, nw + 1), np.float64)
    finish_sums(part, nblk, nw + 1, B, out2)
    assert np.array_equal(out2.get(), got)


@pytest.mark.parametrize("prec", ["fp64", "mixed", "c64"])
@pytest.mark.parametrize("shape", [(64, 256), (128, 1024)])
def test_row_pass_stores_weighted_partials_of_its_output(prec, shape):
    from accuratewigner_amd.device import empty, to_device
    from accuratewigner_amd.engine import FWD, INV, Plan, finish_sums
    rows, cols = shape
    B, nw = 2, 3
    rng = np.random.default_rng(3)
    p = Plan(rows, cols, B, prec)
    z = _rand(rng, (B, rows, cols)).astype(p.state_dtype)
    w = to_device(_rand(rng, (nw, rows, cols)))
    lf = to_device((np.exp(1j * rng.random(rows)) / cols).astype(p.factor_dtype))
    ef = to_device(np.exp(1j * rng.random(cols)).astype(p.factor_dtype))
    a, b = to_device(z), to_device(z)
    nblk = p.row_obs_blocks()
    part = empty((B, nblk, nw), np.float64)
    p.run(a, 1, FWD, INV, line_fac=lf, elem_fac=ef, obs_weights=w, obs_nw=nw, obs_partials=part)
    p.run(b, 1, FWD, INV, line_fac=lf, elem_fac=ef)
    assert np.array_equal(a.get(), b.get())          # the epilogue leaves the output alone
    got = empty((B, nw), np.float64)
    finish_sums(part, nblk, nw, B, got)
    ref = empty((B, nw), np.float64)
    p.weighted_sums(b, w, nw, ref)
    r = ref.get()
    # c64 passes add their FP64 values, the reference reads them back rounded to complex64
    assert np.max(np.abs(got.get() - r)) < (1e-7 if prec == "c64" else 1e-12) * np.max(np.abs(r))


@pytest.mark.parametrize("prec", ["fp64", "mixed", "c64"])
def test_cyclic_diag_obs(prec):
    from accuratewigner_amd.device import empty, to_device
    from accuratewigner_amd.engine import Plan
    n = 256
    rng = np.random.default_rng(4)
    p = Plan(n, n, 1, prec)
    z = _rand(rng, (1, n, n)).astype(p.state_dtype)
    xw = rng.standard_normal((3, n))
    pw = rng.standard_normal((2, n))
    pw_hat = np.fft.fft(pw, axis=1)
    out = empty((1, 8), np.float64)
    ds = empty((1, 2), np.float64)
    p.cyclic_diag_obs(to_device(z), to_device(xw), 3, to_device(pw_hat), 2, out, 8, factor=0.5, dscale=ds)
    zz = z[0].astype(np.complex128)
    i = np.arange(n)
    c = np.array([(-1) ** d * np.sum(zz[i, (i + d) % n]) for d in range(n)])
    diag = zz[i, i].real
    ref = np.concatenate([[diag.sum()], xw @ diag, (pw @ np.fft.fft(c)).real / n])
    got = out.get()[0, :6]
    assert np.max(np.abs(got - ref)) < _tol(prec) * np.max(np.abs(ref))
    assert abs(ds.get()[0, 0] - 1.0 / (got[0] * 0.5)) < 1e-14 * abs(ds.get()[0, 0])


@pytest.mark.parametrize("prec", ["fp64", "mixed", "c64"])
def test_pair_dot(prec):
    from accuratewigner_amd.device import empty, to_device
    from accuratewigner_amd.engine import Plan
    rows, cols, B = 64, 128, 2
    rng = np.random.default_rng(5)
    p = Plan(rows, cols, B, prec)
    a = _rand(rng, (B, rows, cols)).astype(p.state_dtype)
    b = _rand(rng, (B, rows, cols)).astype(p.state_dtype)
    u, v = rng.standard_normal(cols), rng.standard_normal(rows)
    out = empty((B, 4), np.float64)
    p.pair_dot(to_device(a), to_device(b), to_device(u), to_device(v), out)
    ab = np.conj(a.astype(np.complex128)) * b.astype(np.complex128)
    su = np.einsum("bij,j->b", ab, u)
    sv = np.einsum("bij,i->b", ab, v)
    ref = np.stack([su.real, su.imag, sv.real, sv.imag], axis=1)
    assert np.max(np.abs(out.get() - ref)) < _tol(prec) * np.max(np.abs(ref))


@pytest.mark.parametrize("prec", ["fp64", "mixed", "c64"])
@pytest.mark.parametrize("shape", [(256, 128), (1024, 1024)])
def test_column_pass_stores_partials_after_the_factor(prec, shape):
    """obs_mode 1: a [forward, factor] column pass stores the partials of its output v (the
    values right after the factor multiply), weighted sums and sum |v|^2 w2; a [forward,
    factor, inverse] pass stores the same partials and outputs the inverse FFT of v."""
    from accuratewigner_amd.device import empty, to_device
    from accuratewigner_amd.engine import FWD, INV, Plan, finish_sums
    rows, cols = shape
    B, nw = 2, 3
    rng = np.random.default_rng(6)
    p = Plan(rows, cols, B, prec)
    nblk = p.obs_blocks(0)
    assert nblk > 0
    z = _rand(rng, (B, rows, cols)).astype(p.state_dtype)
    w = to_device(_rand(rng, (nw, rows, cols)))
    w2h = rng.random((rows, cols)) + 0.5
    w2 = to_device(w2h)
    lf = to_device(np.exp(1j * rng.random(cols)).astype(p.factor_dtype))   # one per column (line)
    ef = to_device((np.exp(1j * rng.random(rows)) / rows).astype(p.factor_dtype))
    a, b = to_device(z), to_device(z)
    pa, pb = empty((B, nblk, nw + 1), np.float64), empty((B, nblk, nw + 1), np.float64)
    kw = dict(line_fac=lf, elem_fac=ef, obs_weights=w, obs_nw=nw, obs_mode=1, obs_norm2=1, obs_norm2_weights=w2)
    p.run(a, 0, FWD, 0, obs_partials=pa, **kw)
    p.run(b, 0, FWD, INV, obs_partials=pb, **kw)
    ga, gb = empty((B, nw + 1), np.float64), empty((B, nw + 1), np.float64)
    finish_sums(pa, nblk, nw + 1, B, ga)
    finish_sums(pb, nblk, nw + 1, B, gb)
    v = a.get().astype(np.complex128)
    wh = w.get()
    ref = np.concatenate([np.einsum("qij,bij->bq", np.conj(wh), v).real,
                          np.einsum("bij,ij->b", np.abs(v) ** 2, w2h)[:, None]], axis=1)
    # the partials add exactly the stored values in FP64: only the order of the sums differs
    # (c64 passes add their FP64 values before the store rounds them to complex64)
    tol = 1e-7 if prec == "c64" else 1e-12
    assert np.max(np.abs(ga.get() - ref)) < tol * np.max(np.abs(ref))
    assert np.max(np.abs(gb.get() - ga.get())) < 1e-12 * np.max(np.abs(ref))
    c = to_device(a.get())
    p.run(c, 0, 0, INV)
    assert np.max(np.abs(b.get() - c.get())) <= 1e-12 * np.max(np.abs(c.get())) + (0 if prec == "fp64" else 1e-6)


def test_pass_partials_refused_where_blocks_are_not_whole_waves():
    from accuratewigner_amd import _lib
    from accuratewigner_amd.device import empty, to_device
    from accuratewigner_amd.engine import FWD, Plan
    p = Plan(64, 64, 1, "fp64")               # single-state 64-point columns: 2 x 16 threads
    assert p.obs_blocks(0) == 0 and p.obs_blocks(1) > 0
    z = to_device(np.ones((1, 64, 64), np.complex128))
    w = to_device(np.ones((1, 64, 64), np.complex128))
    part = empty((1, 64, 2), np.float64)
    with pytest.raises(_lib.awigInvalidValue):
        p.run(z, 0, FWD, 0, obs_weights=w, obs_nw=1, obs_partials=part, obs_mode=1, obs_norm2=1)
    with pytest.raises(_lib.awigInvalidValue):     # partials of the output: row passes only
        p.run(z, 0, FWD, 0, obs_weights=w, obs_nw=1, obs_partials=part)
    # argument checks, before any launch
    from accuratewigner_amd.engine import INV
    w2 = to_device(np.ones((64, 64)))
    for kw in (dict(obs_nw=1, obs_mode=2),                   # no such mode
               dict(obs_nw=8, obs_norm2=1),                  # more than 8 components
               dict(obs_nw=1, obs_norm2=2),
               dict(obs_nw=1, obs_norm2_weights=w2)):        # weights without the component
        with pytest.raises(_lib.awigInvalidValue):
            p.run(z, 1, FWD, INV, obs_weights=w, obs_partials=part, **kw)


@pytest.mark.parametrize("prec", ["fp64", "mixed", "c64"])
def test_pass_in_scaled_out_equals_scale_pass(prec):
    """A general pass that also stores its input times dscale (the division of rho by its
    norm on the first pass of the Wigner transform) leaves the same bits as a scale pass, and
    its own output equals the pass without the extra store."""
    from accuratewigner_amd.device import to_device
    from accuratewigner_amd.engine import FWD, INV, Plan
    rows, cols, B = 128, 256, 3
    rng = np.random.default_rng(8)
    p = Plan(rows, cols, B, prec)
    z = _rand(rng, (B, rows, cols)).astype(p.state_dtype)
    ds = to_device(np.stack([rng.uniform(0.5, 2.0, B), rng.uniform(-0.5, 0.5, B)], axis=1))
    kw = dict(chirp=0.3, lstep=0.01, lcenter=64., estep=0.02, ecenter=128., sign_in=1, dscale=ds)
    rho, W = to_device(z), to_device(np.zeros_like(z))
    p.run(rho, 1, FWD, INV, out=W, in_scaled_out=rho, **kw)
    rho2, W2 = to_device(z), to_device(np.zeros_like(z))
    p.run(rho2, 1, FWD, INV, out=W2, **kw)
    p.scale(rho2, ds)
    # both round the FP64 product once (the compiler may contract either into an FMA)
    a, b = rho.get().astype(np.complex128), rho2.get().astype(np.complex128)
    assert np.max(np.abs(a - b)) <= 1e-15 * np.max(np.abs(b))
    assert np.array_equal(W.get(), W2.get())
    assert not np.array_equal(rho.get(), z)
