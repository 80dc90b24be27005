"""BF16 / FP16 storage: the default form (FP32 arithmetic in the
compile-time kernels, one rounding at the store) and the MFMA form
(radix-16 stages on MFMA with split hi/lo operands, FP32 accumulation).
The inputs are the 16-bit values; the bar is
the FP64 transform of those values rounded once to the storage type: the
result must equal that correctly rounded answer on all but a tiny fraction
of elements (internal error at FP32 level, far below one 16-bit ulp), and
its rel-L2 must sit at the rounding floor."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [256, 512, 1024, 2048, 4096, 8192, 16384]


def run_case(fy, orc, n, batch, dtype_name, direction, form=0):
    import torch
    tdt = torch.bfloat16 if dtype_name == "bf16" else torch.float16
    prec = fy.BF16 if dtype_name == "bf16" else fy.F16
    g = torch.Generator(device="cuda").manual_seed(n)
    x = torch.randn(batch, n, 2, device="cuda", generator=g).to(tdt)
    with fy.Plan(n, batch, prec, form) as p:
        assert ("mfma16" if form == fy.FORM_MFMA else "lowp-fixed") in p.describe(), p.describe()
        y = p.execute(x, None, direction)
    torch.cuda.synchronize()
    xq = x.float().cpu().numpy().astype(np.float64)
    xc = xq[..., 0] + 1j * xq[..., 1]
    exact = orc.fft_rows(xc, inverse=direction != fy.FORWARD)
    if direction == fy.ADJOINT:
        exact = exact * n
    ex = torch.from_numpy(np.stack([exact.real, exact.imag], -1))
    rounded = ex.to(tdt).double()
    yd = y.cpu().double()
    err = float((yd - ex).norm() / ex.norm())
    floor = float((rounded - ex).norm() / ex.norm())
    mismatch = float((yd != rounded).double().mean())
    return err, floor, mismatch


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("dtype_name", ["bf16", "fp16"])
@pytest.mark.parametrize("form", ["auto", "mfma"])
def test_lowprec_forward_at_rounding_floor(fy, orc, cuda, n, dtype_name, form):
    err, floor, mismatch = run_case(fy, orc, n, 3, dtype_name, fy.FORWARD,
                                    fy.FORM_MFMA if form == "mfma" else fy.FORM_AUTO)
    print(f"N={n} {dtype_name}: rel-L2 {err:.3e}, rounding floor {floor:.3e}, "
          f"elements off the correctly rounded value {mismatch:.2e}")
    assert err < 1.02 * floor + 1e-7
    assert mismatch < (2e-3 if dtype_name == "bf16" else 1e-2)


@pytest.mark.parametrize("dtype_name", ["bf16", "fp16"])
@pytest.mark.parametrize("form", ["auto", "mfma"])
def test_lowprec_inverse_and_adjoint(fy, orc, cuda, dtype_name, form):
    for d in (fy.INVERSE, fy.ADJOINT):
        err, floor, mismatch = run_case(fy, orc, 4096, 2, dtype_name, d,
                                        fy.FORM_MFMA if form == "mfma" else fy.FORM_AUTO)
        assert err < 1.02 * floor + 1e-7, (d, err, floor)


def test_lowprec_unsupported_length(fy, cuda):
    with pytest.raises(fy.FyfftError):
        fy.Plan(6720, 1, fy.BF16)


def test_mfma_form_needs_16bit_storage(fy, cuda):
    with pytest.raises(fy.FyfftError):
        fy.Plan(4096, 1, fy.F32, fy.FORM_MFMA)
