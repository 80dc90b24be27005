"""GPU parity of the four-step (N beyond one workgroup's LDS), Rader and
Bluestein paths against the FP64 oracle, through the C ABI.

Tolerances (relative L2 vs the oracle): FP32 1e-7*log2(N) + 1e-7,
FP64 4e-16*log2(N) + 1e-15 (the convolution paths chain three FFTs)."""
import math
import re

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def tol(n, prec):
    l2 = max(1.0, math.log2(n))
    return 1e-7 * l2 + 1e-7 if prec == "f32" else 4e-16 * l2 + 1e-15


def unit_phase(rng, batch, n, prec):
    """unit-modulus random-phase inputs e^{i phi}, phi ~ U[0, 2pi)"""
    x = np.exp(1j * rng.uniform(0, 2 * np.pi, (batch, n)))
    return x.astype(np.complex64 if prec == "f32" else np.complex128)


def gauss(rng, batch, n, prec):
    x = rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))
    return x.astype(np.complex64 if prec == "f32" else np.complex128)


def check(fy, orc, x, expect_kind, inplace=False):
    import torch
    batch, n = x.shape
    prec = "f32" if x.dtype == np.complex64 else "f64"
    xd = torch.from_numpy(x).cuda()
    with fy.Plan(n, batch, fy.F32 if prec == "f32" else fy.F64) as p:
        desc = p.describe()
        assert desc.startswith(expect_kind), desc
        if inplace:
            y = xd.clone()
            p.execute(y, y, fy.FORWARD)
        else:
            y = p.execute(xd, None, fy.FORWARD)
        z = p.execute(y, None, fy.INVERSE)
        a = p.execute(xd, None, fy.ADJOINT)
    torch.cuda.synchronize()
    x64 = x.astype(np.complex128)
    ref = orc.fft_rows(x64)
    err = orc.rel_l2(y.cpu().numpy(), ref)
    rt = orc.rel_l2(z.cpu().numpy(), x64)
    adj = orc.rel_l2(a.cpu().numpy(), orc.fft_rows(x64, inverse=True) * n)
    print(f"N={n} {prec} {desc}: fwd {err:.3e} round-trip {rt:.3e} adjoint {adj:.3e}")
    t = tol(n, prec)
    assert err < t and adj < t and rt < 2 * t, (err, rt, adj, t)


@pytest.mark.parametrize("n,batch", [(32768, 3), (65536, 2), (78125, 2), (59049, 2),
                                     (1 << 20, 2), (20000 * 7, 1)])
def test_fourstep_f32(fy, orc, cuda, n, batch):
    check(fy, orc, gauss(np.random.default_rng(n), batch, n, "f32"), "fourstep")


@pytest.mark.parametrize("n,batch", [(16384, 2), (65536, 2), (15625, 3), (16807, 2), (19683, 2),
                                     (59049, 1), (78125, 1)])  # odd: compile-time odd-radix factors
def test_fourstep_f64(fy, orc, cuda, n, batch):
    check(fy, orc, gauss(np.random.default_rng(n), batch, n, "f64"), "fourstep")


# Smooth lengths beyond LDS whose runtime-pass sub-transforms take the
# composite radices (two levels in registers): between them every composite
# radix appears, in the column and in the row pass, forward, inverse and
# adjoint.  (n, split, composite radices the plan must show)
COMPOSITE_CASES = [
    (20580, "140(20x7) x 147(21x7)", (20, 21)),
    (21000, "125(5x5x5) x 168(21x8)", (21,)),  # 125: the compile-time odd-radix column kernel
    (21504, "112(16x7) x 192(16x12)", (12,)),
    (21600, "144(16x9) x 150(25x6)", (25, 6)),
    (21870, "135(15x9) x 162(9x9x2)", (15,)),
    (21952, "112(16x7) x 196(14x14)", (14,)),
    (25000, "125(5x5x5) x 200(20x10)", (20, 10)),
    (96000, "240(20x12) x 400(20x20)", (20, 12)),
]


@pytest.mark.parametrize("n,split,radices", COMPOSITE_CASES)
def test_fourstep_composite_radices(fy, orc, cuda, n, split, radices):
    with fy.Plan(n, 1, fy.F32) as p:
        desc = p.describe()
    assert f"N={n} = {split} " in desc, desc
    for r in radices:
        assert re.search(rf"[(x]{r}[x)]", desc), (r, desc)
    check(fy, orc, gauss(np.random.default_rng(n), 2, n, "f32"), "fourstep")
    check(fy, orc, unit_phase(np.random.default_rng(n + 1), 3, n, "f32"), "fourstep", inplace=True)


# Lengths whose two-pass tiles would hold fewer than 8 sub-transforms run
# three passes of short sub-transforms (n = a x b x c)
@pytest.mark.parametrize("n,batch,prec", [(1 << 23, 2, "f32"), (1 << 24, 1, "f32"), (3 ** 15, 1, "f32"),
                                          (1 << 21, 2, "f64"), (1 << 22, 1, "f64")])
def test_fourstep_three_pass(fy, orc, cuda, n, batch, prec):
    with fy.Plan(n, batch, fy.F32 if prec == "f32" else fy.F64) as p:
        assert "three-pass" in p.describe(), p.describe()
    x = gauss(np.random.default_rng(n), batch, n, prec)
    check(fy, orc, x, "fourstep")
    check(fy, orc, x[:1], "fourstep", inplace=True)


def test_fourstep_inplace(fy, orc, cuda):
    check(fy, orc, gauss(np.random.default_rng(1), 2, 65536, "f32"), "fourstep", inplace=True)


@pytest.mark.parametrize("n", [11, 13, 17, 97, 257, 641, 12289])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_rader_lds(fy, orc, cuda, n, prec):
    # FP64 L beyond one workgroup's LDS takes the four-step form
    kind = "rader-lds" if prec == "f32" or n < 8000 else "rader-fourstep"
    check(fy, orc, unit_phase(np.random.default_rng(n), 5, n, prec), kind)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_rader_fourstep_65537(fy, orc, cuda, prec):
    x = unit_phase(np.random.default_rng(65537), 3, 65537, prec)
    check(fy, orc, x, "rader-fourstep")
    check(fy, orc, x, "rader-fourstep", inplace=True)


# four-step Rader through the bucketed permutation: L = N-1 split as
# 140 x 150 / 168 x 180 / 147 x 168 leaves a partial last column group, and
# L is not a multiple of the 4096-element permutation chunk
@pytest.mark.parametrize("n", [21001, 30241, 24697])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_rader_fourstep_partial_groups(fy, orc, cuda, n, prec):
    with fy.Plan(n, 1, fy.F32 if prec == "f32" else fy.F64) as p:
        assert "bucketed-permutation" in p.describe(), p.describe()
    x = gauss(np.random.default_rng(n), 3, n, prec)
    check(fy, orc, x, "rader-fourstep")
    check(fy, orc, x, "rader-fourstep", inplace=True)


@pytest.mark.parametrize("n", [22, 23, 121, 1031, 10006, 10007])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_bluestein_lds(fy, orc, cuda, n, prec):
    if prec == "f64" and n > 5000:
        pytest.skip("covered by the four-step form in FP64")
    check(fy, orc, unit_phase(np.random.default_rng(n), 4, n, prec), "bluestein-lds")


# one prime per compile-time convolution length (M in brackets): 256, 512,
# 1024, 1280, 2048, 2560, 4096, 5120, 8192, 10240, 16384, 20480
# FP64 has compile-time convolutions up to M = 10240 (N = 4099); from
# M = 5120 on they stage the gather through LDS
@pytest.mark.parametrize("n", [103, 223, 419, 521, 821, 1031, 1657, 2053, 3299, 4099, 6563, 8209])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_bluestein_lds_fixed_lengths(fy, orc, cuda, n, prec):
    if prec == "f64" and n > 4099:
        pytest.skip("FP64 compile-time convolutions stop at M = 10240")
    with fy.Plan(n, 5, fy.F32 if prec == "f32" else fy.F64) as p:
        assert "fixed" in p.describe(), p.describe()
    x = unit_phase(np.random.default_rng(n), 5, n, prec)
    check(fy, orc, x, "bluestein-lds")
    check(fy, orc, x, "bluestein-lds", inplace=True)


@pytest.mark.parametrize("n,batch,prec", [(1000003, 2, "f32"), (100003, 2, "f32"),
                                          (10007, 2, "f64")])
def test_bluestein_fourstep(fy, orc, cuda, n, batch, prec):
    x = unit_phase(np.random.default_rng(n), batch, n, prec)
    check(fy, orc, x, "bluestein-fourstep")


def test_bluestein_inplace(fy, orc, cuda):
    x = unit_phase(np.random.default_rng(5), 3, 10007, "f32")
    check(fy, orc, x, "bluestein-lds", inplace=True)


def test_chunked_workspace(fy, orc, cuda):
    """A four-step batch larger than the 1 GiB workspace cap runs in chunks;
    spot-check the first and last transforms."""
    import torch
    n, batch = 1 << 20, 160   # 1.25 GiB of complex64 -> two chunks
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    x = torch.view_as_complex(torch.randn(batch, n, 2, device="cuda", generator=g))
    with fy.Plan(n, batch, fy.F32) as p:
        assert "chunk=128" in p.describe(), p.describe()
        y = p.execute(x, None, fy.FORWARD)
    torch.cuda.synchronize()
    for r in (0, batch - 1):
        xr = x[r].cpu().numpy().astype(np.complex128)
        err = orc.rel_l2(y[r].cpu().numpy(), np.fft.fft(xr))
        assert err < tol(n, "f32"), (r, err)
