"""Strided batches and multi-dimensional transforms against the FP64 oracle
applied along each axis (the multi-dimensional DFT is the product of the
1-D transforms over the axes).  Covers the three pass methods: contiguous
rows, the LDS column kernel, and gather/transform/scatter."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def tol(prec, n):
    return 3e-7 * max(1.0, math.log2(max(n, 2))) if prec == "f32" else 1e-13 * max(1.0, math.log2(max(n, 2)))


def oracle_nd(orc, a, axes, inverse=False):
    out = np.asarray(a, np.complex128)
    for ax in axes:
        moved = np.moveaxis(out, ax, -1)
        shp = moved.shape
        moved = orc.fft_rows(moved.reshape(-1, shp[-1]), inverse=inverse).reshape(shp)
        out = np.moveaxis(moved, -1, ax)
    return out


ND_CASES = [
    ((64, 48), 1), ((1000, 24), 2), ((256, 4096), 1), ((7, 3), 3), ((20000, 3), 1), ((65536, 2), 1),
    ((16, 32, 64), 2), ((5, 7, 9), 1), ((3, 1, 4), 2), ((1000,), 3), ((1, 1), 2), ((12, 10007), 1),
    ((8192, 64), 2), ((6000, 4, 5), 1),   # long first axes: two column passes (colfour)
]


@pytest.mark.parametrize("dims,batch", ND_CASES)
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_nd_vs_oracle(fy, orc, cuda, dims, batch, prec):
    import torch
    cdt = torch.complex64 if prec == "f32" else torch.complex128
    print(f"[nd] dims={dims} batch={batch} prec={prec}", flush=True)
    rng = np.random.default_rng(sum(dims) + batch)
    shape = (batch,) + dims
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    p = fy.NdPlan(dims, batch, fy.F32 if prec == "f32" else fy.F64)
    xd = torch.tensor(x, dtype=cdt, device="cuda")
    y = p.execute(xd)
    torch.cuda.synchronize()
    axes = list(range(1, len(shape)))
    ref = oracle_nd(orc, x, axes)
    n = int(np.prod(dims))
    err = orc.rel_l2(y.cpu().numpy().reshape(shape), ref)
    assert err < tol(prec, n), (p.describe(), err)
    # inverse, in place, against the oracle inverse of the spectrum
    p.execute(y, y, fy.INVERSE)
    torch.cuda.synchronize()
    err = orc.rel_l2(y.cpu().numpy().reshape(shape), x)
    assert err < 2 * tol(prec, n), (p.describe(), err)
    p.close()


def test_nd_methods_and_helpers(fy, orc, cuda):
    import torch
    d = fy.NdPlan((256, 4096), 1, fy.F32).describe()
    assert "rows" in d and "columns" in d, d
    d = fy.NdPlan((65536, 2), 1, fy.F32).describe()
    assert "gather" in d, d
    d = fy.NdPlan((8192, 64), 2, fy.F32).describe()
    assert "colfour" in d, d
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 3, 40, 50)) + 1j * rng.standard_normal((2, 3, 40, 50))
    xd = torch.tensor(x, dtype=torch.complex64, device="cuda")
    y = fy.fft2(xd)
    ref = oracle_nd(orc, x, [2, 3])
    assert orc.rel_l2(y.cpu().numpy(), ref) < 3e-6
    z = fy.ifft2(y)
    assert orc.rel_l2(z.cpu().numpy(), x) < 3e-6
    y3 = fy.fftn(xd, 3)
    assert orc.rel_l2(y3.cpu().numpy(), oracle_nd(orc, x, [1, 2, 3])) < 3e-6


def strided_ref(orc, buf, n, howmany, istride, idist, ostride, odist, out_len, inverse=False):
    rows = np.stack([buf[b * idist + np.arange(n) * istride] for b in range(howmany)])
    res = orc.fft_rows(rows, inverse=inverse)
    out = np.zeros(out_len, np.complex128)
    mask = np.zeros(out_len, bool)
    for b in range(howmany):
        idx = b * odist + np.arange(n) * ostride
        out[idx] = res[b]
        mask[idx] = True
    return out, mask


MANY_CASES = [
    # n, howmany, istride, idist, ostride, odist, method
    (1000, 5, 1, 1024, 1, 1024, "rows"),
    (4096, 6, 1, 4100, 1, 4100, "rows"),
    (65536, 3, 1, 70000, 1, 70000, "rows"),
    (10007, 3, 1, 10010, 1, 10010, "rows"),
    (49, 300, 1, 52, 1, 52, "rows"),        # short length, rows apart: not the staged block copy
    (1000, 5, 1, 1024, 1, 1030, "gather"),
    (100, 7, 7, 1, 7, 1, "columns"),
    (256, 40, 40, 1, 48, 1, "columns"),
    (50, 4, 3, 200, 2, 150, "gather"),
    (6720, 2, 2, 1, 2, 1, "gather"),
    (600, 64, 64, 1, 64, 1, "columns"),
    (8192, 32, 32, 1, 32, 1, "colfour"),    # long adjacent columns: two column passes
    (6000, 20, 20, 1, 24, 1, "colfour"),    # 75 x 80, output pitch differs
    # rows apart from lengths whose batched plan takes a staged compile-time
    # kernel (243) or the lean small-radix kernel (105): the runtime-pass
    # kernel must then run on the compile-time kernel's geometry
    (243, 5, 1, 256, 1, 256, "rows"),
    (105, 9, 1, 112, 1, 112, "rows"),
    (360, 6, 1, 368, 1, 368, "rows"),
    (729, 4, 1, 800, 1, 800, "rows"),
]


@pytest.mark.parametrize("n,howmany,istride,idist,ostride,odist,method", MANY_CASES)
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_many_vs_oracle(fy, orc, cuda, n, howmany, istride, idist, ostride, odist, method, prec):
    import torch
    cdt = torch.complex64 if prec == "f32" else torch.complex128
    print(f"[many] n={n} howmany={howmany} in=({istride},{idist}) out=({ostride},{odist}) {prec}", flush=True)
    p = fy.ManyPlan(n, howmany, istride, idist, ostride, odist, fy.F32 if prec == "f32" else fy.F64)
    assert method in p.describe(), p.describe()
    rng = np.random.default_rng(n + howmany)
    buf = rng.standard_normal(p.in_elems) + 1j * rng.standard_normal(p.in_elems)
    xd = torch.tensor(buf, dtype=cdt, device="cuda")
    for direction in (fy.FORWARD, fy.INVERSE):
        yd = torch.zeros(p.out_elems, dtype=cdt, device="cuda")
        p.execute(xd, yd, direction)
        torch.cuda.synchronize()
        ref, mask = strided_ref(orc, buf, n, howmany, istride, idist, ostride, odist, p.out_elems,
                                inverse=direction != fy.FORWARD)
        y = yd.cpu().numpy()
        err = orc.rel_l2(y[mask], ref[mask])
        assert err < tol(prec, n), (p.describe(), err)
        assert np.all(y[~mask] == 0), "writes outside the output layout"
    # host staging path
    if prec == "f64":
        yh = p.execute_host(buf, np.zeros(p.out_elems, np.complex128))
        ref, mask = strided_ref(orc, buf, n, howmany, istride, idist, ostride, odist, p.out_elems)
        assert orc.rel_l2(yh[mask], ref[mask]) < tol(prec, n)
    p.close()
