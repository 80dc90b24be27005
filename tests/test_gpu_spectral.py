"""Differentiation: the C2C FFT with its hand-written adjoint, and the fused
spectral convolution, checked against FP64 finite differences
(torch.autograd.gradcheck) and against an analytic FP64 reference built from
torch.fft on the CPU."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


# even L with k <= L/2 + 1 take the half-length forms (compile-time passes
# for L/2 = 256 .. 8192, runtime passes for 6720/2); (1024, 1024) and odd L
# take the full-length form; k = L/2 + 1 keeps the Nyquist bin
@pytest.mark.parametrize("L,k", [(256, 32), (4096, 512), (6720, 100), (1024, 1024), (256, 129),
                                 (4096, 2049), (243, 50), (2048, 1), (16384, 700)])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_modes_vs_numpy(fy, cuda, L, k, prec):
    import torch
    rdt = torch.float32 if prec == "f32" else torch.float64
    rng = np.random.default_rng(L + k)
    x = rng.standard_normal((3, 5, L))
    xd = torch.tensor(x, dtype=rdt, device="cuda")
    X = fy.rfft_modes(xd, k, 0.5, 2.0)
    ref = np.fft.fft(x, axis=-1)[..., :k]
    ref[..., 0] *= 0.5
    ref[..., 1:] *= 2.0
    tol = 3e-7 * math.log2(L) if prec == "f32" else 2e-15
    err = np.linalg.norm(X.cpu().numpy() - ref) / np.linalg.norm(ref)
    assert err < tol, err
    m = rng.standard_normal((3, 5, k)) + 1j * rng.standard_normal((3, 5, k))
    md = torch.tensor(m, dtype=torch.complex64 if prec == "f32" else torch.complex128, device="cuda")
    y = fy.irfft_modes(md, L, 1.0 / L, 2.0 / L)
    full = np.zeros((3, 5, L), complex)
    full[..., :k] = m
    full[..., 1:k] *= 2.0
    yref = np.real(np.fft.ifft(full, axis=-1))
    err = np.linalg.norm(y.cpu().numpy() - yref) / np.linalg.norm(yref)
    assert err < tol, err


def test_fft_adjoint_gradcheck(fy, cuda):
    import torch
    from fyfft_amd import autograd as fa
    g = torch.Generator(device="cuda").manual_seed(0)
    for n in (16, 12, 11, 22):
        x = torch.randn(3, n, dtype=torch.complex128, device="cuda", generator=g, requires_grad=True)
        assert torch.autograd.gradcheck(fa.fft, (x,), eps=1e-6, atol=1e-8, rtol=1e-6)
        assert torch.autograd.gradcheck(fa.ifft, (x,), eps=1e-6, atol=1e-8, rtol=1e-6)


def test_adjoint_identity(fy, cuda):
    """<F x, y> = <x, F^H y> with F^H the adjoint direction, FP64."""
    import torch
    g = torch.Generator(device="cuda").manual_seed(1)
    for n in (4096, 6720, 10007, 65537):
        x = torch.randn(2, n, dtype=torch.complex128, device="cuda", generator=g)
        y = torch.randn(2, n, dtype=torch.complex128, device="cuda", generator=g)
        lhs = torch.vdot(fy.fft(x).flatten(), y.flatten())
        rhs = torch.vdot(x.flatten(), fy.adjoint(y).flatten())
        assert abs(lhs - rhs) / abs(lhs) < 1e-13, n


def _reference_layer(x, w):
    """FP64 analytic reference on the CPU: irfft(pad(rfft(x)[:K] mixed by W))."""
    import torch
    L, K = x.shape[-1], w.shape[-1]
    X = torch.fft.rfft(x, dim=-1)[..., :K]
    Y = torch.einsum("bik,iok->bok", X, w)
    Yf = torch.zeros(Y.shape[0], Y.shape[1], L // 2 + 1, dtype=Y.dtype)
    Yf[..., :K] = Y
    return torch.fft.irfft(Yf, n=L, dim=-1)


def test_spectral_conv_gradcheck_fp64(fy, cuda):
    """FP64 finite differences on a [2, 4, 256] slice (4 -> 4 channels, 32 modes)."""
    import torch
    from fyfft_amd.autograd import SpectralConv1dFunction
    g = torch.Generator(device="cuda").manual_seed(2)
    x = torch.randn(2, 4, 256, dtype=torch.float64, device="cuda", generator=g, requires_grad=True)
    w = (torch.randn(4, 4, 32, dtype=torch.complex128, device="cuda", generator=g) / 4).requires_grad_()
    assert torch.autograd.gradcheck(SpectralConv1dFunction.apply, (x, w), eps=1e-6, atol=1e-7, rtol=1e-5)


@pytest.mark.parametrize("shape,modes", [((2, 4, 256), 32), ((8, 64, 4096), 512)])
def test_spectral_conv_vs_analytic(fy, cuda, shape, modes):
    import torch
    from fyfft_amd.autograd import SpectralConv1dFunction
    B, C, L = shape
    rng = np.random.default_rng(B * C)
    x = rng.standard_normal(shape)
    w = (rng.standard_normal((C, C, modes)) + 1j * rng.standard_normal((C, C, modes))) / math.sqrt(2 * C)
    gy = rng.standard_normal(shape)
    # ours, FP32 on the GPU
    xd = torch.tensor(x, dtype=torch.float32, device="cuda", requires_grad=True)
    wd = torch.tensor(w, dtype=torch.complex64, device="cuda", requires_grad=True)
    y = SpectralConv1dFunction.apply(xd, wd)
    y.backward(torch.tensor(gy, dtype=torch.float32, device="cuda"))
    # reference, FP64 on the CPU
    xr = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    wr = torch.tensor(w, dtype=torch.complex128, requires_grad=True)
    yr = _reference_layer(xr, wr)
    yr.backward(torch.tensor(gy, dtype=torch.float64))

    def rel(a, b):
        a = a.detach().cpu().to(b.dtype)
        return float((a - b.detach()).norm() / b.detach().norm())
    ey, ex, ew = rel(y, yr), rel(xd.grad, xr.grad), rel(wd.grad, wr.grad)
    print(f"spectral conv {shape} K={modes}: y {ey:.2e} dx {ex:.2e} dW {ew:.2e}")
    assert ey < 2e-6 and ex < 2e-6 and ew < 2e-6, (ey, ex, ew)


def test_spectral_conv_module_matches_function(fy, cuda):
    """The layer (weights stored mode-major, `weight` a [Ci, Co, K] view)
    gives the function's output and gradients; the weight gradient lands on
    the stored parameter in its own layout."""
    import torch
    from fyfft_amd.autograd import SpectralConv1d, SpectralConv1dFunction
    g = torch.Generator(device="cuda").manual_seed(11)
    layer = SpectralConv1d(8, 6, 40, device="cuda", generator=g)
    assert layer.weight.shape == (8, 6, 40) and layer.weight_kio.shape == (40, 8, 6)
    x = torch.randn(3, 8, 512, device="cuda", generator=g, requires_grad=True)
    gy = torch.randn(3, 6, 512, device="cuda", generator=g)
    y = layer(x)
    y.backward(gy)
    w2 = layer.weight.detach().clone().requires_grad_()
    x2 = x.detach().clone().requires_grad_()
    y2 = SpectralConv1dFunction.apply(x2, w2)
    y2.backward(gy)
    assert torch.allclose(y, y2, rtol=1e-5, atol=1e-6)
    assert torch.allclose(x.grad, x2.grad, rtol=1e-5, atol=1e-6)
    assert torch.allclose(layer.weight_kio.grad.permute(1, 2, 0), w2.grad, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("L,k", [(4096, 512), (6720, 100), (243, 50), (1024, 1024)])
def test_modes_mode_major_layout(fy, cuda, L, k):
    """Mode-major output/input equals the row-major one transposed, for the
    half-length and the full-length forms."""
    import torch
    rng = np.random.default_rng(L * 3 + k)
    x = torch.tensor(rng.standard_normal((4, 6, L)), dtype=torch.float32, device="cuda")
    rm = fy.rfft_modes(x, k, 0.5, 2.0)
    mm = fy.rfft_modes(x, k, 0.5, 2.0, mode_major=True)
    assert mm.shape == (k, 4, 6)
    assert torch.equal(mm.permute(1, 2, 0), rm)
    m = torch.tensor(rng.standard_normal((4, 6, k)) + 1j * rng.standard_normal((4, 6, k)),
                     dtype=torch.complex64, device="cuda")
    yr = fy.irfft_modes(m, L, 1.0 / L, 2.0 / L)
    ym = fy.irfft_modes(m.permute(2, 0, 1).contiguous(), L, 1.0 / L, 2.0 / L, mode_major=True)
    assert ym.shape == (4, 6, L)
    assert torch.equal(ym, yr)


@pytest.mark.parametrize("K,R,ci,co", [(7, 37, 19, 23), (3, 1, 1, 1), (5, 130, 64, 70), (512, 256, 64, 64)])
@pytest.mark.parametrize("dtype", ["c64", "c128"])
def test_mix_modes_vs_fp64(fy, cuda, K, R, ci, co, dtype):
    """The MFMA mode-mixing kernel (FP32 and FP64 matrix cores) against an
    FP64 batched matmul on the CPU, all three forms, shapes that are and are
    not multiples of the 64x64x16 tiling.  Tolerance: 2e-6 (complex64),
    1e-14 (complex128)."""
    import torch
    rng = np.random.default_rng(K * 1000 + R)

    def cplx(*shape):
        return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    xm, w, gy = cplx(K, R, ci), cplx(K, ci, co), cplx(K, R, co)
    d = lambda a: torch.tensor(a, dtype=torch.complex64 if dtype == "c64" else torch.complex128, device="cuda")
    t = lambda a: torch.tensor(a, dtype=torch.complex128)
    cases = [(fy.MIX_FORWARD, xm, w, t(xm) @ t(w)),
             (fy.MIX_ADJOINT_X, gy, w, t(gy) @ t(w).mH),
             (fy.MIX_GRAD_W, xm, gy, t(xm).mH @ t(gy))]
    for kind, p, q, ref in cases:
        out = fy.mix_modes(kind, d(p), d(q))
        torch.cuda.synchronize()
        err = float((out.cpu().to(torch.complex128) - ref).norm() / ref.norm())
        print(f"mix kind {kind} K={K} R={R} ci={ci} co={co}: rel {err:.2e}")
        assert out.shape == ref.shape and err < (2e-6 if dtype == "c64" else 1e-14), (kind, err)
