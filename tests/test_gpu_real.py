"""Real-input (R2C) and real-output (C2R) plans against the FP64 oracle.

R2C is checked against the oracle's complex transform of the real rows
(bins 0..n/2); C2R against the oracle's inverse of the Hermitian spectrum
rebuilt from a random half spectrum, and by the round trip C2R(R2C(x)) = x.
The lengths cover every plan form: even n through a half-length transform
fused in LDS or run as a four-step plus a post/pre pass, odd n fused at full
length, and lengths whose half (or whole) length needs Rader / Bluestein
through the full-length passes."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    # (n, rows, form)
    (2, 7, "fused full"), (4, 5, "fused half"), (16, 9, "fused half"), (1000, 6, "fused half"),
    (1024, 5, "fused half"), (4096, 4, "fused half"), (6720, 3, "fused half"), (20000, 2, "fused half"),
    (1, 3, "full-length transform"), (3, 5, "fused full"), (15, 4, "fused full"), (243, 3, "fused full"),
    (1125, 3, "fused full"), (16807, 2, "fused full"),
    (65536, 3, "half-length transform"), (1 << 21, 1, "half-length transform"),
    (10007, 2, "full-length transform"), (20014, 2, "full-length transform"),
    (177147, 1, "full-length transform"),
]


def tol(prec, n):
    return 3e-7 * max(1.0, math.log2(max(n, 2))) if prec == "f32" else 1e-13 * max(1.0, math.log2(max(n, 2)))


def half_spectrum_ref(orc, x):
    return orc.fft_rows(x.astype(np.complex128))[:, : x.shape[1] // 2 + 1]


@pytest.mark.parametrize("n,rows,form", CASES)
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_r2c_c2r_vs_oracle(fy, orc, cuda, n, rows, form, prec):
    import torch
    if prec == "f64" and n >= (1 << 21):
        pytest.skip("covered in f32; keeps the f64 oracle run short")
    rdt = torch.float32 if prec == "f32" else torch.float64
    cdt = torch.complex64 if prec == "f32" else torch.complex128
    print(f"[real] n={n} rows={rows} prec={prec}", flush=True)
    rng = np.random.default_rng(n * 7 + rows)
    x = rng.standard_normal((rows, n))
    p = fy.RealPlan(n, rows, fy.F32 if prec == "f32" else fy.F64)
    if prec == "f32" and form.startswith("fused"):
        assert form in p.describe(), p.describe()
    xd = torch.tensor(x, dtype=rdt, device="cuda")
    X = p.execute(xd)
    torch.cuda.synchronize()
    ref = half_spectrum_ref(orc, x)
    err = orc.rel_l2(X.cpu().numpy(), ref)
    assert err < tol(prec, n), (p.describe(), err)
    # round trip
    y = p.execute(X, None, fy.INVERSE)
    torch.cuda.synchronize()
    err = orc.rel_l2(y.cpu().numpy(), x)
    assert err < 2 * tol(prec, n), (p.describe(), err)
    # C2R of an arbitrary half spectrum against the oracle inverse
    h = n // 2 + 1
    H = rng.standard_normal((rows, h)) + 1j * rng.standard_normal((rows, h))
    full = np.zeros((rows, n), np.complex128)
    full[:, :h] = H
    # a real signal's bins 0 and n/2 are real: C2R ignores their imaginary parts
    full[:, 0] = full[:, 0].real
    if n % 2 == 0:
        full[:, n // 2] = full[:, n // 2].real
    if n > 1:
        full[:, h:] = np.conj(H[:, 1: n - h + 1][:, ::-1])
    yref = orc.fft_rows(full, inverse=True).real
    y = p.execute(torch.tensor(H, dtype=cdt, device="cuda"), None, fy.INVERSE)
    torch.cuda.synchronize()
    err = orc.rel_l2(y.cpu().numpy(), yref)
    assert err < tol(prec, n), (p.describe(), err)
    p.close()


def test_real_helpers_and_errors(fy, orc, cuda):
    import torch
    rng = np.random.default_rng(5)
    x = rng.standard_normal((4, 3, 960))
    xd = torch.tensor(x, dtype=torch.float32, device="cuda")
    X = fy.rfft(xd)
    assert X.shape == (4, 3, 481) and X.dtype == torch.complex64
    ref = orc.fft_rows(x.reshape(-1, 960).astype(np.complex128))[:, :481].reshape(4, 3, 481)
    assert orc.rel_l2(X.cpu().numpy(), ref) < 3e-6
    y = fy.irfft(X, 960)
    assert y.shape == (4, 3, 960)
    assert orc.rel_l2(y.cpu().numpy(), x) < 3e-6
    # host staging path
    p = fy.RealPlan(960, 12, fy.F64)
    Xh = p.execute_host(x.reshape(12, 960))
    assert orc.rel_l2(Xh, ref.reshape(12, 481)) < 1e-13
    yh = p.execute_host(Xh, None, fy.INVERSE)
    assert orc.rel_l2(yh, x.reshape(12, 960)) < 1e-13
    # the adjoint direction is not defined for real plans
    with pytest.raises(fy.FyfftError):
        p.execute_ptr(xd.data_ptr(), X.data_ptr(), fy.ADJOINT)
    p.close()
