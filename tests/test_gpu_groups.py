"""Step sequences over batches larger than 192 MiB run in groups of items (so that a group
stays in the MALL between passes).  The grouped run must be bit-identical to running the
same items a few at a time, below the grouping threshold, with the per-item factor tables,
masks and opening scales offset by hand."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 1024


def _unit(rng, shape, scale=1.0):
    return scale * np.exp(2j * np.pi * rng.random(shape))


@pytest.mark.parametrize("prec,B,per_call", [("mixed", 27, 3), ("fp64", 15, 3), ("c64", 27, 3)])
def test_grouped_sequence_is_bit_identical(prec, B, per_call):
    from accuratewigner_amd.device import to_device
    from accuratewigner_amd.engine import Plan
    n = 4
    plan = Plan(N, N, B, prec)
    cdt, fdt = plan.state_dtype, plan.factor_dtype
    esz = np.dtype(cdt).itemsize
    fsz = np.dtype(fdt).itemsize     # factor tables and masks follow the arithmetic precision
    rsz = fsz // 2
    item = N * N * esz
    # 27 x 8 MiB (mixed) and 15 x 16 MiB (fp64) exceed the 192 MiB group size: two groups;
    # 3 items per call (24 / 48 MiB, 3072 lines) run as one launch per pass
    assert B * item > 192 << 20 and per_call * item < 192 << 20
    rng = np.random.default_rng(5)
    W0 = (rng.standard_normal((B, N, N)) + 1j * rng.standard_normal((B, N, N))).astype(cdt)
    # per-item tables: column vectors v = 0..n, row vectors s = 0..n-1 (1/N scales keep O(1))
    cl = to_device(_unit(rng, (B, n + 1, N)).astype(fdt))
    ce = to_device(_unit(rng, (B, n + 1, N), 1.0 / N).astype(fdt))
    rl = to_device(_unit(rng, (B, n, N), 1.0 / N).astype(fdt))
    re = to_device(_unit(rng, (B, n, N)).astype(fdt))
    bs_c, bs_r = (n + 1) * N, n * N
    # per-item masks (one table for the open/close, merged and row passes each)
    masks = [to_device((0.5 + 0.5 * rng.random((B, N, N))).astype(plan.real_dtype)) for _ in range(3)]
    ds = to_device(np.stack([rng.uniform(0.5, 1.5, B), rng.uniform(-0.5, 0.5, B)], axis=1))
    kw = dict(col_line_bstride=bs_c, col_elem_bstride=bs_c, row_line_bstride=bs_r, row_elem_bstride=bs_r,
              mask_bstride=N * N)

    a = to_device(W0)
    plan.steps(a, n, cl, ce, rl, re, mask_open_close=masks[0], mask_mid=masks[1], mask_row=masks[2],
               dscale=ds, **kw)

    b = to_device(W0)
    for g in range(0, B, per_call):
        plan.steps(b.ptr + g * item, n, cl.ptr + g * bs_c * fsz, ce.ptr + g * bs_c * fsz,
                   rl.ptr + g * bs_r * fsz, re.ptr + g * bs_r * fsz,
                   mask_open_close=masks[0].ptr + g * N * N * rsz, mask_mid=masks[1].ptr + g * N * N * rsz,
                   mask_row=masks[2].ptr + g * N * N * rsz, dscale=ds.ptr + g * 16, batch=per_call, **kw)
    ga, gb = a.get(), b.get()
    assert np.all(np.isfinite(ga))
    assert np.array_equal(ga, gb)
    # the sequence did change every item
    assert np.min(np.linalg.norm((ga - W0).reshape(B, -1), axis=1)) > 0
