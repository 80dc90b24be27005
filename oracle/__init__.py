"""CPU oracle for the 1-D complex FFT path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg import this module, and only as the checker (or, for
``cpu_baseline``, as the timed CPU port of the reference algorithm).  The
product package ``fyfft_amd`` never imports it.

The C restatement lives in ``oracle/oracle.c`` + ``oracle/oracle_impl.inc``;
``oracle/Makefile`` builds ``oracle/liboracle.so``.  Each function there cites
the reference file:line it follows.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

KERNEL_DIRECT, MIX_RADIX, RADER, BLUESTEIN, DFT = range(5)
STRATEGY_NAMES = ["kernel_direct", "mix_radix", "rader", "bluestein", "DFT"]


def build() -> str:
    """Compile liboracle.so from the C sources (gcc via oracle/Makefile)."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = ctypes.CDLL(_LIB_PATH)
    sz, i, dp, fp = ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p
    for name in ("fyo_is_directly_supported", "fyo_mix_is_processable", "fyo_is_prime",
                 "fyo_rader_is_processable", "fyo_bluestein_is_processable", "fyo_strategy"):
        getattr(L, name).argtypes = [sz]
        getattr(L, name).restype = i
    for name in ("fyo_next_smooth", "fyo_compute_n1", "fyo_primitive_root", "fyo_rader_M",
                 "fyo_bluestein_M"):
        getattr(L, name).argtypes = [sz]
        getattr(L, name).restype = sz
    L.fyo_smooth_length.argtypes = [i]
    L.fyo_smooth_length.restype = sz
    L.fyo_power_mod.argtypes = [sz, sz, sz]
    L.fyo_power_mod.restype = sz
    L.fyo_decomposition.argtypes = [sz, ctypes.POINTER(sz)]
    L.fyo_decomposition.restype = i
    for name in ("fyo_dft_naive_f64", "fyo_fft_f64", "fyo_fft_ext_f64", "fyo_fft_f32"):
        getattr(L, name).argtypes = [sz, dp, dp, i]
        getattr(L, name).restype = i
    L.fyo_fft_f32_batch_timed.argtypes = [sz, sz, fp, fp, i, i]
    L.fyo_fft_f32_batch_timed.restype = ctypes.c_double
    _lib = L
    return L


# ---------------------------------------------------------------- planner
def is_directly_supported(n: int) -> bool:
    return bool(lib().fyo_is_directly_supported(n))


def smooth_lengths() -> list[int]:
    return [lib().fyo_smooth_length(i) for i in range(38)]


def next_smooth(n: int) -> int:
    return lib().fyo_next_smooth(n)


def compute_n1(n: int) -> int:
    return lib().fyo_compute_n1(n)


def mix_is_processable(n: int) -> bool:
    return bool(lib().fyo_mix_is_processable(n))


def decomposition(n: int) -> list[int]:
    arr = (ctypes.c_size_t * 10)()
    cnt = lib().fyo_decomposition(n, arr)
    if cnt < 0:
        raise ValueError(f"no decomposition for {n}")
    return [arr[k] for k in range(cnt)]


def is_prime(n: int) -> bool:
    return bool(lib().fyo_is_prime(n))


def power_mod(b: int, e: int, m: int) -> int:
    return lib().fyo_power_mod(b, e, m)


def primitive_root(n: int) -> int:
    return lib().fyo_primitive_root(n)


def rader_is_processable(n: int) -> bool:
    return bool(lib().fyo_rader_is_processable(n))


def rader_M(n: int) -> int:
    return lib().fyo_rader_M(n)


def bluestein_is_processable(n: int) -> bool:
    return bool(lib().fyo_bluestein_is_processable(n))


def bluestein_M(n: int) -> int:
    return lib().fyo_bluestein_M(n)


def strategy(n: int) -> int:
    return lib().fyo_strategy(n)


# ---------------------------------------------------------------- transforms
def _call(fn, x: np.ndarray, dtype, inverse: bool) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=dtype)
    if x.ndim != 1:
        raise ValueError("1-D input expected")
    n = x.shape[0]
    out = np.empty_like(x)
    st = fn(n, x.ctypes.data, out.ctypes.data, int(inverse))
    if st != 0:
        raise RuntimeError(f"oracle transform failed for N={n} (status {st})")
    return out


def dft_naive(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    """FP64 O(N^2) DFT with exact n*k mod N reduction (the FP64 bar)."""
    return _call(lib().fyo_dft_naive_f64, x, np.complex128, inverse)


def fft64(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    """The reference's strategy ladder restated in double."""
    return _call(lib().fyo_fft_f64, x, np.complex128, inverse)


def fft32(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    """The reference's strategy ladder restated in float (the CPU port)."""
    return _call(lib().fyo_fft_f32, x, np.complex64, inverse)


def fft_ext(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    """FP64 FFT for any N (ladder where the reference is feasible,
    power-of-two Bluestein elsewhere)."""
    return _call(lib().fyo_fft_ext_f64, x, np.complex128, inverse)


def fft_rows(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    """FP64 oracle applied to each row of a 2-D array."""
    x = np.asarray(x)
    return np.stack([fft_ext(r, inverse) for r in x.reshape(-1, x.shape[-1])]).reshape(x.shape)


def fft32_batch_timed(x: np.ndarray, inverse: bool = False, threads: int = 1):
    """Time the FP32 port over the rows of x on `threads` threads.
    Returns (output, seconds)."""
    x = np.ascontiguousarray(x, dtype=np.complex64)
    batch, n = x.shape
    out = np.empty_like(x)
    sec = lib().fyo_fft_f32_batch_timed(n, batch, x.ctypes.data, out.ctypes.data,
                                        int(inverse), int(threads))
    if sec < 0:
        raise RuntimeError("oracle batch port failed")
    return out, sec


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))
