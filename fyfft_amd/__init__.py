"""fyfft_amd -- MI355X-native batched 1-D complex FFT.

Python binding over the C ABI of ``fyfft_amd/lib/libfyfft.so``
(declared in ``include/fyfft.h``).  PyTorch is used only for device memory
and streams; every transform runs in the library's HIP kernels.  There is no
CPU fallback: if the library is missing, loading raises.

Mirror of the reference interface (kjear/fyFFT, ``foye_FFT.hpp``):

* :class:`FFT1D_C2C_invoker` -- ``FFT1D_C2C_invoker(N)`` +
  ``dispatch(input, output, invert)`` on host (numpy) buffers with the
  reference's synchronous semantics (``foye_FFT.hpp:101-151``), and
  ``acquired_execution_strategy(N)`` (``foye_FFT.hpp:162-184``).
* :class:`Plan` -- the device-pointer fast path: batched, stream-ordered.
"""
from __future__ import annotations

import ctypes
import enum
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libfyfft.so")

min_input_length = 2          # foye_FFT.hpp:86
max_input_length = 80_000     # foye_FFT.hpp:87 (the reference's limit)
MAX_LENGTH = 1 << 24          # this library's limit (FYFFT_MAX_LENGTH)

FORWARD, INVERSE, ADJOINT = 0, 1, 2
F32, F64, BF16, F16 = 0, 1, 2, 3
FORM_AUTO, FORM_PLAIN, FORM_TUNED, FORM_MFMA = 0, 1, 2, 3


class _Options(ctypes.Structure):
    _fields_ = [("form", ctypes.c_int), ("reserved", ctypes.c_int * 7)]


class FFT_execution_strategy(enum.IntEnum):
    """Same order as the reference enum (foye_FFT.hpp:153-160)."""
    kernel_direct = 0
    mix_radix = 1
    rader = 2
    bluestein = 3
    DFT = 4


_STATUS = {0: "success", 1: "invalid argument", 2: "invalid length", 3: "unsupported",
           4: "allocation failed", 5: "device error", 6: "invalid plan", 7: "no device"}


class FyfftError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        detail = ""
        try:
            detail = _load().fyfft_last_error().decode()
        except Exception:  # pragma: no cover - library itself failed to load
            pass
        super().__init__(f"{what}: {_STATUS.get(status, status)}" + (f" ({detail})" if detail else ""))


_lib = None
_lib_lock = threading.Lock()


def _load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        # torch ships its own libamdhip64.so.7; loading it first makes this
        # library bind to the same HIP runtime instead of a second copy
        # from /opt/rocm (two runtimes in one process do not share devices).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"fyfft_amd: native library not built ({LIB_PATH}); run "
                "`python __graft_entry__.py` or `python fyfft_amd/build.py`")
        L = ctypes.CDLL(LIB_PATH)
        vp, sz, i, ip = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_int)
        L.fyfft_plan_create.argtypes = [ctypes.POINTER(vp), sz, sz, i]
        L.fyfft_plan_destroy.argtypes = [vp]
        L.fyfft_execute.argtypes = [vp, vp, vp, i, vp]
        L.fyfft_execute_host.argtypes = [vp, vp, vp, i]
        L.fyfft_execute_noalloc.argtypes = [vp, vp, vp, i, vp]
        L.fyfft_plan_strategy.argtypes = [vp, ip]
        L.fyfft_length_strategy.argtypes = [sz, ip]
        L.fyfft_length_radices.argtypes = [sz, ip, i, ip]
        L.fyfft_plan_workspace_bytes.argtypes = [vp, ctypes.POINTER(sz)]
        L.fyfft_plan_describe.argtypes = [vp, ctypes.c_char_p, sz]
        dbl = ctypes.c_double
        L.fyfft_rfft_modes.argtypes = [vp, vp, vp, sz, dbl, dbl, vp]
        L.fyfft_irfft_modes.argtypes = [vp, vp, vp, sz, dbl, dbl, vp]
        L.fyfft_rfft_modes_ex.argtypes = [vp, vp, vp, sz, dbl, dbl, i, vp]
        L.fyfft_irfft_modes_ex.argtypes = [vp, vp, vp, sz, dbl, dbl, i, vp]
        L.fyfft_mix_modes.argtypes = [i, i, vp, vp, vp, sz, sz, sz, sz, vp]
        L.fyfft_plan_create_grouped.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(sz), sz, i, ctypes.c_uint]
        L.fyfft_plan_elements.argtypes = [vp, ctypes.POINTER(sz)]
        L.fyfft_plan_create_ex.argtypes = [ctypes.POINTER(vp), sz, sz, i, ctypes.POINTER(_Options)]
        L.fyfft_plan_create_real.argtypes = [ctypes.POINTER(vp), sz, sz, i]
        L.fyfft_plan_create_many.argtypes = [ctypes.POINTER(vp), sz, sz, ctypes.POINTER(_Layout), i]
        L.fyfft_plan_create_nd.argtypes = [ctypes.POINTER(vp), i, ctypes.POINTER(sz), sz, i]
        L.fyfft_wisdom_export.argtypes = [ctypes.c_char_p, sz, ctypes.POINTER(sz)]
        L.fyfft_wisdom_import.argtypes = [ctypes.c_char_p]
        L.fyfft_wisdom_forget.argtypes = []
        L.fyfft_status_string.argtypes = [i]
        L.fyfft_status_string.restype = ctypes.c_char_p
        L.fyfft_last_error.argtypes = []
        L.fyfft_last_error.restype = ctypes.c_char_p
        for name in ("fyfft_plan_create", "fyfft_plan_destroy", "fyfft_execute", "fyfft_execute_noalloc",
                     "fyfft_execute_host", "fyfft_plan_strategy", "fyfft_length_strategy",
                     "fyfft_length_radices", "fyfft_plan_workspace_bytes", "fyfft_plan_describe",
                     "fyfft_rfft_modes", "fyfft_irfft_modes", "fyfft_rfft_modes_ex", "fyfft_irfft_modes_ex",
                     "fyfft_mix_modes", "fyfft_plan_create_grouped",
                     "fyfft_plan_elements", "fyfft_plan_create_ex", "fyfft_plan_create_real",
                     "fyfft_plan_create_many", "fyfft_plan_create_nd",
                     "fyfft_wisdom_export",
                     "fyfft_wisdom_import", "fyfft_wisdom_forget"):
            getattr(L, name).restype = i
        _lib = L
        return L


def lib() -> ctypes.CDLL:
    """The loaded libfyfft.so (raises ImportError if it was never built)."""
    return _load()


def _check(status: int, what: str) -> None:
    if status != 0:
        raise FyfftError(status, what)


# ------------------------------------------------------------------ wisdom
def wisdom_export() -> str:
    """Autotuner choices of this process as text."""
    need = ctypes.c_size_t()
    _check(_load().fyfft_wisdom_export(None, 0, ctypes.byref(need)), "wisdom_export")
    buf = ctypes.create_string_buffer(need.value)
    _check(_load().fyfft_wisdom_export(buf, need.value, None), "wisdom_export")
    return buf.value.decode()


def wisdom_import(text: str) -> None:
    _check(_load().fyfft_wisdom_import(text.encode()), "wisdom_import")


def wisdom_forget() -> None:
    _check(_load().fyfft_wisdom_forget(), "wisdom_forget")


# ------------------------------------------------------------------ queries
def length_strategy(n: int) -> FFT_execution_strategy:
    out = ctypes.c_int()
    _check(_load().fyfft_length_strategy(n, ctypes.byref(out)), f"strategy(N={n})")
    return FFT_execution_strategy(out.value)


def length_radices(n: int) -> list[int]:
    buf = (ctypes.c_int * 32)()
    cnt = ctypes.c_int()
    _check(_load().fyfft_length_radices(n, buf, 32, ctypes.byref(cnt)), f"radices(N={n})")
    return [buf[k] for k in range(cnt.value)]


def _precision_of(dtype) -> int:
    """complex64 -> F32, complex128 -> F64; bfloat16 / float16 tensors whose
    last dimension is (re, im) -> BF16 / F16 storage."""
    import torch
    if dtype == torch.complex64:
        return F32
    if dtype == torch.complex128:
        return F64
    if dtype == torch.bfloat16:
        return BF16
    if dtype == torch.float16 or dtype == torch.complex32:
        return F16
    raise TypeError(f"fyfft_amd: unsupported dtype {dtype}")


def _complex_numel(t) -> int:
    import torch
    if t.dtype in (torch.bfloat16, torch.float16):
        return t.numel() // 2
    return t.numel()


# ------------------------------------------------------------------ plan
class Plan:
    """A batched device plan: ``batch`` contiguous transforms of length ``n``.

    ``execute(x, out, direction)`` takes torch tensors on the plan's device
    whose last dimension holds the transforms; it enqueues on the current
    torch stream (or ``stream``) and does not synchronise.
    """

    def __init__(self, n: int, batch: int = 1, precision: int = F32, form: int = FORM_AUTO):
        self.n, self.batch, self.precision = int(n), int(batch), int(precision)
        self._h = ctypes.c_void_p()
        opts = _Options(int(form))
        _check(_load().fyfft_plan_create_ex(ctypes.byref(self._h), self.n, self.batch, self.precision,
                                            ctypes.byref(opts)),
               f"plan_create(N={n}, batch={batch}, precision={precision}, form={form})")

    def close(self) -> None:
        if self._h:
            _load().fyfft_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def strategy(self) -> FFT_execution_strategy:
        out = ctypes.c_int()
        _check(_load().fyfft_plan_strategy(self._h, ctypes.byref(out)), "plan_strategy")
        return FFT_execution_strategy(out.value)

    @property
    def workspace_bytes(self) -> int:
        out = ctypes.c_size_t()
        _check(_load().fyfft_plan_workspace_bytes(self._h, ctypes.byref(out)), "workspace_bytes")
        return out.value

    def describe(self) -> str:
        buf = ctypes.create_string_buffer(512)
        _check(_load().fyfft_plan_describe(self._h, buf, 512), "describe")
        return buf.value.decode()

    def _check_tensor(self, t, name):
        import torch
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise TypeError(f"{name} must be a device tensor")
        if _precision_of(t.dtype) != self.precision:
            raise TypeError(f"{name} dtype {t.dtype} does not match the plan precision")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous")
        if _complex_numel(t) != self.n * self.batch:
            raise ValueError(f"{name} has {_complex_numel(t)} complex elements, plan needs "
                             f"{self.n * self.batch}")

    def execute_ptr(self, in_ptr: int, out_ptr: int, direction: int = FORWARD, stream: int = 0) -> None:
        _check(_load().fyfft_execute(self._h, ctypes.c_void_p(in_ptr), ctypes.c_void_p(out_ptr),
                                     int(direction), ctypes.c_void_p(stream)), "execute")

    def execute(self, x, out=None, direction: int = FORWARD, stream=None):
        import torch
        self._check_tensor(x, "input")
        if out is None:
            out = torch.empty_like(x)
        self._check_tensor(out, "output")
        if stream is None:
            stream = torch.cuda.current_stream(x.device)
        handle = stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream)
        self.execute_ptr(x.data_ptr(), out.data_ptr(), direction, handle)
        return out

    def execute_host(self, x: np.ndarray, out: np.ndarray | None = None, direction: int = FORWARD):
        if self.precision not in (F32, F64):
            raise TypeError("execute_host: use device tensors for BF16/F16 plans")
        dt = np.complex64 if self.precision == F32 else np.complex128
        x = np.ascontiguousarray(x, dtype=dt)
        if x.size != self.n * self.batch:
            raise ValueError("host buffer size does not match the plan")
        if out is None:
            out = np.empty_like(x)
        _check(_load().fyfft_execute_host(self._h, ctypes.c_void_p(x.ctypes.data),
                                          ctypes.c_void_p(out.ctypes.data), int(direction)),
               "execute_host")
        return out


def _stream_handle(stream, device):
    import torch
    if stream is None:
        stream = torch.cuda.current_stream(device)
    return stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream)


def rfft_modes(x, k: int, scale0: float = 1.0, scale: float = 1.0, stream=None, mode_major: bool = False):
    """First k modes of the C2C FFT of real rows x [..., L] (float32/float64
    device tensor): out[..., j] = c_j * sum_n x[..., n] e^{-2 pi i n j / L},
    c_0 = scale0, c_j = scale.  Returns complex [..., k], or [k, ...] with
    mode_major (one contiguous [rows] slab per mode, the layout of per-mode
    GEMMs)."""
    import torch
    if x.dtype not in (torch.float32, torch.float64) or not x.is_cuda:
        raise TypeError("rfft_modes: real float32/float64 device tensor expected")
    x = x.contiguous()
    L = x.shape[-1]
    rows = x.numel() // L
    prec = F32 if x.dtype == torch.float32 else F64
    p = _cached_plan(L, rows, prec, x.device.index or 0)
    shape = (k, *x.shape[:-1]) if mode_major else (*x.shape[:-1], k)
    out = torch.empty(shape, dtype=torch.complex64 if prec == F32 else torch.complex128, device=x.device)
    _check(_load().fyfft_rfft_modes_ex(p._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                       k, float(scale0), float(scale), 1 if mode_major else 0,
                                       ctypes.c_void_p(_stream_handle(stream, x.device))), "rfft_modes")
    return out


def irfft_modes(modes, L: int, scale0: float, scale: float, stream=None, mode_major: bool = False):
    """y[..., n] = Re(sum_{j<k} c_j modes[..., j] e^{+2 pi i n j / L}) for
    complex modes [..., k] (or [k, ...] with mode_major); c_0 = scale0,
    c_j = scale.  Returns real [..., L]."""
    import torch
    if modes.dtype not in (torch.complex64, torch.complex128) or not modes.is_cuda:
        raise TypeError("irfft_modes: complex64/complex128 device tensor expected")
    modes = modes.contiguous()
    k = modes.shape[0] if mode_major else modes.shape[-1]
    lead = modes.shape[1:] if mode_major else modes.shape[:-1]
    rows = modes.numel() // k
    prec = _precision_of(modes.dtype)
    p = _cached_plan(L, rows, prec, modes.device.index or 0)
    y = torch.empty(*lead, L, dtype=torch.float32 if prec == F32 else torch.float64, device=modes.device)
    _check(_load().fyfft_irfft_modes_ex(p._h, ctypes.c_void_p(modes.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                                        k, float(scale0), float(scale), 1 if mode_major else 0,
                                        ctypes.c_void_p(_stream_handle(stream, modes.device))), "irfft_modes")
    return y


MIX_FORWARD, MIX_ADJOINT_X, MIX_GRAD_W = 0, 1, 2


def mix_modes(kind: int, p, q, stream=None):
    """Per-mode complex GEMMs of the spectral convolution on the matrix cores
    (complex64 or complex128 device tensors, mode-major): with modes [K, rows, C] and
    weights [K, Ci, Co],
      MIX_FORWARD:   p modes [K, R, Ci], q weights -> [K, R, Co]  (p @ q)
      MIX_ADJOINT_X: p modes [K, R, Co], q weights -> [K, R, Ci]  (p @ q^H)
      MIX_GRAD_W:    p modes [K, R, Ci], q modes [K, R, Co] -> [K, Ci, Co]  (p^H @ q)."""
    import torch
    if p.dtype not in (torch.complex64, torch.complex128) or q.dtype != p.dtype or not p.is_cuda or not q.is_cuda:
        raise TypeError("mix_modes: complex64 or complex128 device tensors of one dtype expected")
    p, q = p.contiguous(), q.contiguous()
    if p.dim() != 3 or q.dim() != 3 or p.shape[0] != q.shape[0]:
        raise ValueError("mix_modes: [K, ., .] operands with equal K expected")
    K = p.shape[0]
    if kind == MIX_FORWARD:
        R, ci, co = p.shape[1], p.shape[2], q.shape[2]
        if q.shape[1] != ci:
            raise ValueError("mix_modes: weights must be [K, Ci, Co]")
        out = torch.empty(K, R, co, dtype=p.dtype, device=p.device)
    elif kind == MIX_ADJOINT_X:
        R, co, ci = p.shape[1], p.shape[2], q.shape[1]
        if q.shape[2] != co:
            raise ValueError("mix_modes: weights must be [K, Ci, Co]")
        out = torch.empty(K, R, ci, dtype=p.dtype, device=p.device)
    elif kind == MIX_GRAD_W:
        R, ci, co = p.shape[1], p.shape[2], q.shape[2]
        if q.shape[1] != R:
            raise ValueError("mix_modes: both operands must be [K, R, .]")
        out = torch.empty(K, ci, co, dtype=p.dtype, device=p.device)
    else:
        raise ValueError("mix_modes: unknown kind")
    prec = F32 if p.dtype == torch.complex64 else F64
    _check(_load().fyfft_mix_modes(kind, prec, ctypes.c_void_p(p.data_ptr()), ctypes.c_void_p(q.data_ptr()),
                                   ctypes.c_void_p(out.data_ptr()), K, R, ci, co,
                                   ctypes.c_void_p(_stream_handle(stream, p.device))), "mix_modes")
    return out


class _Layout(ctypes.Structure):
    _fields_ = [("istride", ctypes.c_size_t), ("idist", ctypes.c_size_t),
                ("ostride", ctypes.c_size_t), ("odist", ctypes.c_size_t)]


class _StridedPlan(Plan):
    """Common execution of strided plans: the tensors are flat device
    buffers covering the input / output layouts."""

    in_elems = out_elems = 0

    def _check_buf(self, t, name, need):
        import torch
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise TypeError(f"{name} must be a device tensor")
        if _precision_of(t.dtype) != self.precision:
            raise TypeError(f"{name} dtype {t.dtype} does not match the plan precision")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous")
        if t.numel() < need:
            raise ValueError(f"{name} has {t.numel()} elements, the layout spans {need}")

    def execute(self, x, out=None, direction: int = FORWARD, stream=None):
        import torch
        self._check_buf(x, "input", self.in_elems)
        if out is None:
            out = torch.zeros(max(self.out_elems, 1), dtype=x.dtype, device=x.device)
            if self.out_elems == x.numel():
                out = out.view(x.shape)
        self._check_buf(out, "output", self.out_elems)
        self.execute_ptr(x.data_ptr(), out.data_ptr(), direction, _stream_handle(stream, x.device))
        return out

    def execute_host(self, x: np.ndarray, out: np.ndarray | None = None, direction: int = FORWARD):
        dt = np.complex64 if self.precision == F32 else np.complex128
        x = np.ascontiguousarray(x, dtype=dt)
        if x.size != self.in_elems:
            raise ValueError(f"host input has {x.size} elements, the layout spans {self.in_elems}")
        out = np.zeros(self.out_elems, dt) if out is None else out
        if out.dtype != dt or out.size != self.out_elems or not out.flags.c_contiguous:
            raise ValueError("host output must be a contiguous buffer spanning the output layout")
        _check(_load().fyfft_execute_host(self._h, ctypes.c_void_p(x.ctypes.data),
                                          ctypes.c_void_p(out.ctypes.data), int(direction)),
               "execute_host")
        return out


class ManyPlan(_StridedPlan):
    """``howmany`` transforms of length ``n``; element j of transform b at
    x[b*idist + j*istride] in the input and y[b*odist + j*ostride] in the
    output (element units, flat buffers)."""

    def __init__(self, n: int, howmany: int, istride: int = 1, idist: int | None = None,
                 ostride: int | None = None, odist: int | None = None, precision: int = F32):
        idist = n * istride if idist is None else idist
        ostride = istride if ostride is None else ostride
        odist = idist if odist is None else odist
        self.n, self.batch, self.precision = int(n), int(howmany), int(precision)
        self.layout = (int(istride), int(idist), int(ostride), int(odist))
        self._h = ctypes.c_void_p()
        lay = _Layout(*self.layout)
        _check(_load().fyfft_plan_create_many(ctypes.byref(self._h), self.n, self.batch, ctypes.byref(lay),
                                              self.precision),
               f"plan_create_many(N={n}, howmany={howmany}, layout={self.layout})")
        self.in_elems = (howmany - 1) * idist + (n - 1) * istride + 1
        self.out_elems = (howmany - 1) * odist + (n - 1) * ostride + 1


class NdPlan(_StridedPlan):
    """``batch`` contiguous row-major arrays of shape ``dims`` (rank 1..3);
    the transform runs over every axis (inverse scaled by 1/prod(dims))."""

    def __init__(self, dims, batch: int = 1, precision: int = F32):
        self.dims = tuple(int(d) for d in dims)
        self.batch, self.precision = int(batch), int(precision)
        self.n = int(np.prod(self.dims))
        arr = (ctypes.c_size_t * len(self.dims))(*self.dims)
        self._h = ctypes.c_void_p()
        _check(_load().fyfft_plan_create_nd(ctypes.byref(self._h), len(self.dims), arr, self.batch,
                                            self.precision),
               f"plan_create_nd(dims={self.dims}, batch={batch})")
        self.in_elems = self.out_elems = self.n * self.batch


_nd_cache: dict = {}


def fftn(x, ndim: int | None = None, inverse: bool = False, stream=None):
    """C2C transform over the last ``ndim`` (1..3, default all up to 3) axes
    of a complex64/complex128 device tensor."""
    import torch
    if x.dtype not in (torch.complex64, torch.complex128) or not x.is_cuda:
        raise TypeError("fftn: complex64/complex128 device tensor expected")
    ndim = min(x.dim(), 3) if ndim is None else int(ndim)
    if not 1 <= ndim <= min(3, x.dim()):
        raise ValueError("fftn: 1..3 transform axes")
    x = x.contiguous()
    dims = tuple(x.shape[x.dim() - ndim:])
    batch = x.numel() // int(np.prod(dims)) if x.numel() else 0
    key = (dims, batch, _precision_of(x.dtype), x.device.index or 0)
    with _plan_cache_lock:
        p = _nd_cache.get(key)
        if p is None:
            with torch.cuda.device(x.device):
                p = NdPlan(dims, batch, _precision_of(x.dtype))
            if len(_nd_cache) > 32:
                _nd_cache.clear()
            _nd_cache[key] = p
    out = torch.empty_like(x)
    p.execute_ptr(x.data_ptr(), out.data_ptr(), INVERSE if inverse else FORWARD, _stream_handle(stream, x.device))
    return out


def ifftn(x, ndim: int | None = None, stream=None):
    return fftn(x, ndim, True, stream)


def fft2(x, stream=None):
    return fftn(x, 2, False, stream)


def ifft2(x, stream=None):
    return fftn(x, 2, True, stream)


class RealPlan(Plan):
    """``batch`` rows of real length ``n``: ``execute(x, X)`` (FORWARD, R2C)
    maps real rows [.., n] to complex [.., n//2 + 1]; ``execute(X, x,
    INVERSE)`` (C2R) maps a Hermitian half spectrum back to real rows,
    scaled by 1/n.  Input and output must not overlap."""

    def __init__(self, n: int, batch: int = 1, precision: int = F32):
        self.n, self.batch, self.precision = int(n), int(batch), int(precision)
        self._h = ctypes.c_void_p()
        _check(_load().fyfft_plan_create_real(ctypes.byref(self._h), self.n, self.batch, self.precision),
               f"plan_create_real(N={n}, batch={batch}, precision={precision})")

    def _check_side(self, t, name, real):
        import torch
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise TypeError(f"{name} must be a device tensor")
        want = ((torch.float32 if self.precision == F32 else torch.float64) if real
                else (torch.complex64 if self.precision == F32 else torch.complex128))
        if t.dtype != want:
            raise TypeError(f"{name} dtype {t.dtype}, plan needs {want}")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous")
        need = self.batch * (self.n if real else self.n // 2 + 1)
        if t.numel() != need:
            raise ValueError(f"{name} has {t.numel()} elements, plan needs {need}")

    def execute(self, x, out=None, direction: int = FORWARD, stream=None):
        import torch
        r2c = direction == FORWARD
        self._check_side(x, "input", r2c)
        if out is None:
            dt = ((torch.complex64 if self.precision == F32 else torch.complex128) if r2c
                  else (torch.float32 if self.precision == F32 else torch.float64))
            shape = tuple(x.shape[:-1]) + ((self.n // 2 + 1) if r2c else self.n,)
            out = torch.empty(shape, dtype=dt, device=x.device)
        self._check_side(out, "output", not r2c)
        self.execute_ptr(x.data_ptr(), out.data_ptr(), direction, _stream_handle(stream, x.device))
        return out

    def execute_host(self, x: np.ndarray, out: np.ndarray | None = None, direction: int = FORWARD):
        r2c = direction == FORWARD
        rdt, cdt = (np.float32, np.complex64) if self.precision == F32 else (np.float64, np.complex128)
        x = np.ascontiguousarray(x, dtype=rdt if r2c else cdt)
        if x.size != self.batch * (self.n if r2c else self.n // 2 + 1):
            raise ValueError("host buffer size does not match the plan")
        if out is None:
            out = np.empty((self.batch, self.n // 2 + 1) if r2c else (self.batch, self.n),
                           dtype=cdt if r2c else rdt)
        _check(_load().fyfft_execute_host(self._h, ctypes.c_void_p(x.ctypes.data),
                                          ctypes.c_void_p(out.ctypes.data), int(direction)),
               "execute_host")
        return out


_real_cache: dict = {}


def _cached_real_plan(n: int, batch: int, precision: int, device: int) -> RealPlan:
    import torch
    key = (n, batch, precision, device)
    with _plan_cache_lock:
        p = _real_cache.get(key)
        if p is None:
            with torch.cuda.device(device):
                p = RealPlan(n, batch, precision)
            if len(_real_cache) > 64:
                _real_cache.clear()
            _real_cache[key] = p
        return p


def rfft(x, stream=None):
    """Real-input FFT over the last dimension of a float32/float64 device
    tensor: complex [..., n//2 + 1], unnormalised (torch.fft.rfft semantics)."""
    import torch
    if x.dtype not in (torch.float32, torch.float64) or not x.is_cuda:
        raise TypeError("rfft: real float32/float64 device tensor expected")
    x = x.contiguous()
    n = x.shape[-1]
    p = _cached_real_plan(n, x.numel() // n, F32 if x.dtype == torch.float32 else F64, x.device.index or 0)
    return p.execute(x, None, FORWARD, stream)


def irfft(X, n: int, stream=None):
    """Inverse of rfft: real [..., n] from the half spectrum [..., n//2 + 1],
    scaled by 1/n (torch.fft.irfft semantics)."""
    import torch
    if X.dtype not in (torch.complex64, torch.complex128) or not X.is_cuda:
        raise TypeError("irfft: complex64/complex128 device tensor expected")
    X = X.contiguous()
    if X.shape[-1] != n // 2 + 1:
        raise ValueError(f"irfft: last dimension {X.shape[-1]} != n//2 + 1 = {n // 2 + 1}")
    p = _cached_real_plan(n, X.numel() // X.shape[-1], _precision_of(X.dtype), X.device.index or 0)
    return p.execute(X, None, INVERSE, stream)


GROUPED_GRAPH = 1
GROUPED_PER_TRANSFORM = 2
GROUPED_TUNED = 4


class GroupedPlan(Plan):
    """Ragged batch: transforms of lengths `lengths`, packed back to back in
    one 1-D device tensor; executed as one grouped launch per phase.
    flags: GROUPED_GRAPH (capture/replay a hipGraph per (in, out, direction,
    stream)), GROUPED_PER_TRANSFORM (baseline: one plain launch per
    transform)."""

    def __init__(self, lengths, precision: int = F32, flags: int = 0):
        lens = [int(n) for n in lengths]
        arr = (ctypes.c_size_t * len(lens))(*lens)
        self.lengths = lens
        self.precision = int(precision)
        self.offsets = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64) if lens else []
        self.total = int(sum(lens))
        self.n, self.batch = self.total, 1
        self._h = ctypes.c_void_p()
        _check(_load().fyfft_plan_create_grouped(ctypes.byref(self._h), arr, len(lens), self.precision,
                                                 int(flags)), f"plan_create_grouped(count={len(lens)})")


_plan_cache: dict = {}
_plan_cache_lock = threading.Lock()


def _cached_plan(n: int, batch: int, precision: int, device: int) -> Plan:
    import torch
    key = (n, batch, precision, device)
    with _plan_cache_lock:
        p = _plan_cache.get(key)
        if p is None:
            with torch.cuda.device(device):
                p = Plan(n, batch, precision)
            if len(_plan_cache) > 64:
                _plan_cache.clear()
            _plan_cache[key] = p
        return p


def _run(x, direction: int):
    import torch
    if not x.is_cuda:
        raise TypeError("fyfft_amd.fft: input must be a device tensor")
    if x.dtype not in (torch.complex64, torch.complex128):
        raise TypeError("fyfft_amd.fft: complex64 / complex128 input expected")
    x = x.contiguous()
    n = x.shape[-1]
    batch = x.numel() // n if n else 0
    p = _cached_plan(n, batch, _precision_of(x.dtype), x.device.index or 0)
    return p.execute(x, torch.empty_like(x), direction)


def fft(x):
    """Forward C2C FFT over the last dimension (unnormalised)."""
    return _run(x, FORWARD)


def ifft(x):
    """Inverse C2C FFT over the last dimension (scaled by 1/N)."""
    return _run(x, INVERSE)


def adjoint(x):
    """Adjoint of the forward transform (= N * ifft)."""
    return _run(x, ADJOINT)


# ------------------------------------------------------------------ drop-in mirror
class FFT1D_C2C_invoker:
    """Host-pointer drop-in for ``fy::fft::FFT1D_C2C_invoker``
    (foye_FFT.hpp:89-185).

    ``dispatch(input, output, invert)`` transforms one length-N sequence of
    interleaved (re, im) float32 values, synchronously, like
    ``dispatch<invert>(const float*, float*)``.  ``input``/``output`` may be
    float32 arrays of 2N values or complex64 arrays of N values; in place is
    allowed.  Out-of-range N raises ``ValueError`` where the reference has
    undefined behaviour (foye_FFT.hpp:103-104).
    """

    FFT_execution_strategy = FFT_execution_strategy

    def __init__(self, length: int):
        n = int(length)
        if not (min_input_length <= n <= MAX_LENGTH):
            raise ValueError(f"FFT1D_C2C_invoker: N={n} outside [{min_input_length}, {MAX_LENGTH}]")
        self.N = n
        self._plan = Plan(n, 1, F32)

    def dispatch(self, input, output, invert: bool = False) -> None:
        src = np.asarray(input)
        dst = output
        if src.dtype == np.float32:
            src = src.view(np.complex64)
        if dst.dtype == np.float32:
            dst_c = dst.view(np.complex64)
        else:
            dst_c = dst
        if src.size != self.N or dst_c.size != self.N:
            raise ValueError("dispatch: buffers must hold N complex values")
        if dst_c.dtype != np.complex64 or not dst_c.flags.c_contiguous:
            raise TypeError("dispatch: output must be contiguous float32/complex64")
        src = np.ascontiguousarray(src, dtype=np.complex64)
        _check(_load().fyfft_execute_host(self._plan._h, ctypes.c_void_p(src.ctypes.data),
                                          ctypes.c_void_p(dst_c.ctypes.data),
                                          INVERSE if invert else FORWARD), "dispatch")

    @staticmethod
    def acquired_execution_strategy(length: int) -> FFT_execution_strategy:
        return length_strategy(int(length))
