"""Grouped ragged transforms: one persistent grouped launch per phase over a
device task table, checked transform by transform against the FP64 oracle,
against the per-transform baseline form, and under hipGraph replay."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL32 = lambda n: 1e-7 * max(1.0, math.log2(n)) + 1e-7


def packed(rng, lengths, dtype=np.complex64):
    tot = int(sum(lengths))
    return (rng.standard_normal(tot) + 1j * rng.standard_normal(tot)).astype(dtype)


def segments(arr, lengths):
    off = 0
    for n in lengths:
        yield arr[off:off + n]
        off += n


# every kind of member: LDS smooth / Rader / Bluestein, four-step smooth,
# Rader and Bluestein, up to the configuration's 2^20 and a 1000003 prime
MIXED = [1, 2, 3, 16, 17, 100, 105, 243, 1024, 4096, 4096, 4096, 6720, 10007, 12289, 16384,
         16807, 23, 65536, 65537, 97, 140000, 100003, 2, 2, 7, 121, 1031, 4096, 30011,
         1 << 20, 1000003]


@pytest.mark.parametrize("direction", ["forward", "inverse", "adjoint"])
def test_grouped_vs_oracle(fy, orc, cuda, direction):
    import torch
    rng = np.random.default_rng(7)
    x = packed(rng, MIXED)
    d = {"forward": fy.FORWARD, "inverse": fy.INVERSE, "adjoint": fy.ADJOINT}[direction]
    with fy.GroupedPlan(MIXED, fy.F32) as p:
        print(p.describe())
        xd = torch.from_numpy(x).cuda()
        y = p.execute(xd, None, d).cpu().numpy()
    torch.cuda.synchronize()
    for n, xs, ys in zip(MIXED, segments(x, MIXED), segments(y, MIXED)):
        ref = orc.fft_ext(xs.astype(np.complex128), inverse=direction != "forward")
        if direction == "adjoint":
            ref = ref * n
        err = orc.rel_l2(ys, ref)
        assert err < TOL32(n), (n, err)


def test_grouped_f64(fy, orc, cuda):
    import torch
    lens = [16, 4096, 6720, 10007, 65537, 30011, 13, 8192]
    rng = np.random.default_rng(8)
    x = packed(rng, lens, np.complex128)
    with fy.GroupedPlan(lens, fy.F64) as p:
        y = p.execute(torch.from_numpy(x).cuda(), None, fy.FORWARD).cpu().numpy()
    for n, xs, ys in zip(lens, segments(x, lens), segments(y, lens)):
        err = orc.rel_l2(ys, orc.fft_ext(xs))
        assert err < 4e-16 * math.log2(n) + 1e-15, (n, err)


def test_grouped_matches_per_transform_and_inplace(fy, cuda):
    import torch
    from fyfft_amd.workloads import ragged_lengths
    lens = ragged_lengths(count=600, seed=3, total_target=600 * 8000.0)
    rng = np.random.default_rng(9)
    x = torch.from_numpy(packed(rng, lens)).cuda()
    with fy.GroupedPlan(lens, fy.F32) as g, \
         fy.GroupedPlan(lens, fy.F32, fy.GROUPED_PER_TRANSFORM) as b:
        yg = g.execute(x, None, fy.FORWARD)
        yb = b.execute(x, None, fy.FORWARD)
        xi = x.clone()
        g.execute(xi, xi, fy.FORWARD)
    torch.cuda.synchronize()
    # same bodies and tables; only the code around them differs
    assert float((yg - yb).norm() / yb.norm()) < 1e-6
    assert torch.equal(yg, xi)


def test_grouped_graph_replay(fy, cuda):
    """hipGraph capture/replay: every replay, with eager work and
    synchronisation in between, equals the eager result."""
    import torch
    from fyfft_amd.workloads import ragged_lengths
    lens = ragged_lengths(count=300, seed=4, total_target=300 * 6000.0)
    rng = np.random.default_rng(10)
    x = torch.from_numpy(packed(rng, lens)).cuda()
    s = torch.cuda.Stream()
    with fy.GroupedPlan(lens, fy.F32) as eager_p, \
         fy.GroupedPlan(lens, fy.F32, fy.GROUPED_GRAPH) as graph_p:
        ref = eager_p.execute(x, None, fy.FORWARD)
        y = torch.empty_like(x)
        for rep in range(4):
            y.zero_()
            torch.cuda.synchronize()
            with torch.cuda.stream(s):
                graph_p.execute(x, y, fy.FORWARD, stream=s)
            s.synchronize()
            assert torch.equal(y, ref), rep
            _ = (x * 2).sum().item()  # unrelated eager work between replays
        z = torch.empty_like(x)
        with torch.cuda.stream(s):
            graph_p.execute(y, z, fy.INVERSE, stream=s)
        s.synchronize()
    assert float((z - x).norm() / x.norm()) < 1e-6


# 16-bit storage: the same members; the bar is the FP64 transform of the
# 16-bit input values, rounded once to the storage type (the internal error
# is at FP32 level, far below one 16-bit ulp)
LOWP = [16, 100, 243, 4096, 6720, 97, 12289, 23, 1031, 65536, 140000, 65537, 100003, 1 << 20, 1000003]


@pytest.mark.parametrize("dtype_name", ["bf16", "fp16"])
@pytest.mark.parametrize("direction", ["forward", "inverse"])
def test_grouped_lowp_at_rounding_floor(fy, orc, cuda, dtype_name, direction):
    import torch
    tdt = torch.bfloat16 if dtype_name == "bf16" else torch.float16
    prec = fy.BF16 if dtype_name == "bf16" else fy.F16
    d = fy.FORWARD if direction == "forward" else fy.INVERSE
    tot = sum(LOWP)
    g = torch.Generator(device="cuda").manual_seed(11)
    x = torch.randn(tot, 2, device="cuda", generator=g).to(tdt)
    print(f"[grouped lowp] {dtype_name} {direction}", flush=True)
    with fy.GroupedPlan(LOWP, prec) as p, fy.GroupedPlan(LOWP, prec, fy.GROUPED_PER_TRANSFORM) as b:
        assert f"storage={dtype_name}" in p.describe(), p.describe()
        y = p.execute(x, None, d)
        yb = b.execute(x, None, d)
    torch.cuda.synchronize()
    xq = x.float().cpu().numpy().astype(np.float64)
    xc = xq[:, 0] + 1j * xq[:, 1]
    yh = y.cpu().double().numpy()
    ybh = yb.cpu().double().numpy()
    off, bad, total = 0, 0, 0
    for n in LOWP:
        ex = orc.fft_ext(xc[off:off + n], inverse=d != fy.FORWARD)
        ext = np.stack([ex.real, ex.imag], -1)
        rounded = torch.from_numpy(ext).to(tdt).double().numpy()
        seg = yh[off:off + n]
        err = np.linalg.norm(seg - ext) / np.linalg.norm(ext)
        floor = np.linalg.norm(rounded - ext) / np.linalg.norm(ext)
        bad += int((seg != rounded).sum())
        total += seg.size
        assert err < 1.02 * floor + 1e-6, (n, err, floor)
        # the baseline (FP32 plain kernels between widening and rounding copies)
        errb = np.linalg.norm(ybh[off:off + n] - ext) / np.linalg.norm(ext)
        assert errb < 1.02 * floor + 1e-6, (n, errb, floor)
        off += n
    assert bad / total < (2e-3 if dtype_name == "bf16" else 1e-2), bad / total
