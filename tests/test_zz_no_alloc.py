"""The allocation-free weighted sums (aw_weighted_sums_noalloc, through the C ABI) against
aw_weighted_sums on a plan: bit-identical, on two calls in a row (the second smaller), in all
three precisions."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(lib, prec_code, rows, cols, batch, data, w, nw, norm2, out, factor, ds):
    from accuratewigner_amd import _lib
    from accuratewigner_amd.device import current_stream
    vp = ctypes.c_void_p
    _lib.check(lib.aw_weighted_sums_noalloc(prec_code, rows, cols, batch, vp(data.ptr), vp(w.ptr), nw, int(norm2),
                                            vp(out.ptr), factor, vp(ds.ptr), vp(current_stream())))


@pytest.mark.parametrize("prec", ["fp64", "mixed", "c64"])
def test_weighted_sums_noalloc_equals_plan_entry_point(prec):
    from accuratewigner_amd import _lib
    from accuratewigner_amd.device import empty, to_device
    from accuratewigner_amd.engine import Plan
    lib = _lib.load()
    rng = np.random.default_rng(21)
    for rows, cols, batch, nw, norm2 in ((256, 512, 5, 3, True), (64, 128, 2, 1, False)):
        p = Plan(rows, cols, batch, prec)
        z = (rng.standard_normal((batch, rows, cols)) + 1j * rng.standard_normal((batch, rows, cols))).astype(p.state_dtype)
        w = to_device(rng.standard_normal((nw, rows, cols)) + 1j * rng.standard_normal((nw, rows, cols)))
        zd = to_device(z)
        nc = nw + int(norm2)
        a, b = empty((batch, nc), np.float64), empty((batch, nc), np.float64)
        da, db = empty((batch, 2), np.float64), empty((batch, 2), np.float64)
        p.weighted_sums(zd, w, nw, a, factor=0.5, dscale=da, norm2=norm2)
        _call(lib, p.prec, rows, cols, batch, zd, w, nw, norm2, b, 0.5, db)
        assert np.array_equal(a.get(), b.get()), (rows, cols)
        assert np.array_equal(da.get(), db.get())
        ref = np.einsum("qij,bij->bq", np.conj(w.get()), z.astype(np.complex128)).real
        assert np.max(np.abs(b.get()[:, :nw] - ref)) < 1e-10 * np.max(np.abs(ref))


def test_weighted_sums_noalloc_rejects_oversized_input():
    from accuratewigner_amd import _lib
    lib = _lib.load()
    vp = ctypes.c_void_p
    # 512 blocks x 8 components x 257 items exceeds the fixed scratch
    rc = lib.aw_weighted_sums_noalloc(0, 1024, 1024, 257, vp(1), vp(1), 7, 1, vp(1), 0.0, vp(None), vp(None))
    assert rc == 8 and b"scratch" in lib.aw_last_error()
