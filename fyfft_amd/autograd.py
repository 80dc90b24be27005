"""Differentiable transforms over the C ABI.

* :func:`fft` / :func:`ifft` -- C2C transforms whose backward is the
  hand-written adjoint kernel: for y = F x the gradient is F^H g
  (FYFFT_ADJOINT: conjugate twiddles, no 1/N); for y = F^{-1} x = F^H x / N it
  is F g / N.
* :class:`SpectralConv1d` -- the 1-D spectral convolution
      y = irfft( (rfft(x)[..., :K]) mixed over channels by W )
  computed as two fused transforms (real-input FFT keeping K modes, K-mode
  real inverse) around a batched complex GEMM over the modes; its backward
  is written by hand from the same kernels.

Gradients follow torch's convention for complex tensors
(dL/dRe z + i dL/dIm z).
"""
from __future__ import annotations

import torch

from . import (ADJOINT, FORWARD, INVERSE, MIX_ADJOINT_X, MIX_FORWARD, MIX_GRAD_W, _run, irfft_modes,
               mix_modes, rfft_modes)


class _FFT(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return _run(x, FORWARD)

    @staticmethod
    def backward(ctx, g):
        return _run(g.contiguous(), ADJOINT)


class _IFFT(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return _run(x, INVERSE)

    @staticmethod
    def backward(ctx, g):
        n = g.shape[-1]
        return _run(g.contiguous(), FORWARD) / n


def fft(x):
    """Differentiable forward C2C FFT over the last dimension."""
    return _FFT.apply(x)


def ifft(x):
    """Differentiable inverse C2C FFT (1/N) over the last dimension."""
    return _IFFT.apply(x)


class SpectralConv1dFunction(torch.autograd.Function):
    """y = Re(IFFT_L(Z)), Z[0] = Y[0], Z[k] = 2 Y[k] (0 < k < K), Z = 0 above,
    Y[b,o,k] = sum_i X[b,i,k] W[i,o,k], X = FFT_L(x)[..., :K]."""

    # The modes live mode-major ([K, B, C]) between the transforms, so each
    # mode's channel mixing is one GEMM with no layout copies of the
    # [B, C, K] mode tensors: the MFMA kernel of fyfft_mix_modes (forward,
    # p @ q^H and p^H @ q forms; FP32 MFMA for complex64, FP64 MFMA for
    # complex128).  W [Ci, Co, K] is
    # used through its [K, Ci, Co] permutation, a view without a copy when W
    # is stored mode-major (as SpectralConv1d stores it), and dW is returned
    # as the same permuted view of a [K, Ci, Co] result.

    @staticmethod
    def forward(ctx, x, w):
        L = x.shape[-1]
        K = w.shape[-1]
        if 2 * K > L:
            raise ValueError("SpectralConv1d: modes must satisfy 2*K <= L")
        xm = rfft_modes(x, K, mode_major=True)               # [K, B, Ci]
        wt = w.permute(2, 0, 1)                               # [K, Ci, Co]
        if not wt.is_contiguous():
            wt = wt.contiguous()
        ym = mix_modes(MIX_FORWARD, xm, wt)
        y = irfft_modes(ym, L, 1.0 / L, 2.0 / L, mode_major=True)   # [B, Co, L]
        ctx.save_for_backward(xm, wt)
        ctx.L = L
        return y

    @staticmethod
    def backward(ctx, g):
        xm, wt = ctx.saved_tensors
        L, K = ctx.L, wt.shape[0]
        # dL/dY = S * (1/L) * FFT(g) restricted to the kept modes (S = 1, 2, 2, ...)
        gy = rfft_modes(g.contiguous(), K, 1.0 / L, 2.0 / L, mode_major=True)   # [K, B, Co]
        gxm = mix_modes(MIX_ADJOINT_X, gy, wt)                                  # [K, B, Ci]
        gw = mix_modes(MIX_GRAD_W, xm, gy).permute(1, 2, 0)                     # [Ci, Co, K] view
        # dL/dx = Re(F^H pad(gX)): the real part of the adjoint over K modes
        gx = irfft_modes(gxm, L, 1.0, 1.0, mode_major=True)
        return gx, gw


class SpectralConv1d(torch.nn.Module):
    """1-D spectral convolution layer: in_channels -> out_channels keeping
    `modes` Fourier modes; weights ~ N(0, 1/in_channels) complex.

    ``weight`` is the [in, out, modes] complex tensor; it is stored mode-major
    (``weight_kio``, [modes, in, out]) so the per-mode GEMMs use it and produce
    its gradient without layout copies, and ``weight`` is a view of it."""

    def __init__(self, in_channels: int, out_channels: int, modes: int, dtype=torch.complex64,
                 device=None, generator=None):
        super().__init__()
        scale = (1.0 / in_channels) ** 0.5
        real_dt = torch.float32 if dtype == torch.complex64 else torch.float64
        w = torch.randn(in_channels, out_channels, modes, 2, dtype=real_dt, device=device,
                        generator=generator) * (scale / 2 ** 0.5)
        self.weight_kio = torch.nn.Parameter(torch.view_as_complex(w).permute(2, 0, 1).contiguous())

    @property
    def weight(self):
        return self.weight_kio.permute(1, 2, 0)

    def forward(self, x):
        return SpectralConv1dFunction.apply(x, self.weight)
