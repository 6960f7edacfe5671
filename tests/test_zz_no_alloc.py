"""The allocation-free matvec (mpsd_heff_apply_noalloc) against the allocating entry point
mpsd_heff_apply on the same tensors, through the C-ABI: bit-identical results (the same kernels
in the same order) and the oracle's product to 1e-12, on two calls in a row, the second with a
smaller input; no device memory is added to the context; shapes beyond the static buffer are
rejected."""
import numpy as np
import pytest

from oracle import mpsd_oracle as orc

pytestmark = pytest.mark.gpu


def _operands(D, Dr, w, d, seed):
    L = orc.complex_gaussian(seed + 1, (D, w, D))
    W = orc.complex_gaussian(seed + 2, (d, w, d, w))
    R = orc.complex_gaussian(seed + 3, (Dr, w, Dr))
    x = orc.complex_gaussian(seed + 4, (d * D * Dr,))
    return L, W, R, x


def test_noalloc_matches_heff_apply():
    from mpsdyn_amd import device as dev
    ctx = dev.Context(0)
    # (D, Dr, w, d, offset): the split last stage at D = 200, then a smaller problem, then one
    # that takes the unsplit last stage (d = 3)
    cases = [(200, 200, 3, 2, 0.7), (40, 56, 5, 2, 0.0), (24, 24, 4, 3, -1.25)]
    ops = []
    for k, (D, Dr, w, d, off) in enumerate(cases):
        L, W, R, x = _operands(D, Dr, w, d, seed=10 * k)
        ops.append((L, W, R, x, off, ctx.tensor(L), ctx.tensor(W), ctx.tensor(R), ctx.tensor(x),
                    ctx.tensor(dims=(d * D * Dr,))))
    ctx.sync()
    before = ctx.device_bytes()
    got = []
    for (L, W, R, x, off, tL, tW, tR, tx, ty) in ops:
        print(f"noalloc D={L.shape[0]} Dr={R.shape[0]} w={W.shape[1]} d={W.shape[0]}", flush=True)
        ctx.check(ctx.lib.mpsd_heff_apply_noalloc(ctx.handle, tL.handle, tW.handle, tR.handle, off, tx.devptr(),
                                                  ty.devptr()))
        got.append(ty.download())
    ctx.sync()
    assert ctx.device_bytes() == before  # nothing allocated by the calls
    for (L, W, R, x, off, tL, tW, tR, tx, ty), g in zip(ops, got):
        h = dev.Heff(ctx, tL, tW, tR, offset=off)
        ref = ctx.tensor(dims=(g.size,))
        h.apply_device(tx.devptr(), ref.devptr())
        want = ref.download()
        h.close()
        assert np.array_equal(g, want)
        o = orc.heff_product(L, W, R, x, offset=off)
        assert np.linalg.norm(g - o) <= 1e-12 * np.linalg.norm(o)
    ctx.close()


def test_noalloc_rejects_shapes_beyond_the_buffer():
    from mpsdyn_amd import device as dev
    ctx = dev.Context(0)
    D, w = 512, 5  # T1 + T2 = 2 * 2 * 5 * 512^2 complex = 84 MB > 64 MiB
    tL = ctx.tensor(dims=(D, w, D))
    tW = ctx.tensor(dims=(2, w, 2, w))
    tx = ctx.tensor(dims=(2 * D * D,))
    ty = ctx.tensor(dims=(2 * D * D,))
    rc = ctx.lib.mpsd_heff_apply_noalloc(ctx.handle, tL.handle, tW.handle, tL.handle, 0.0, tx.devptr(), ty.devptr())
    assert rc == -6
    assert b"static buffer" in ctx.lib.mpsd_last_error(ctx.handle)
    ctx.close()
