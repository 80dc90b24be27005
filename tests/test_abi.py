"""libfyfft.so loads and exports every symbol include/fyfft.h declares;
the planner queries and argument checks work without a GPU (no compute
calls here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "fyfft.h")).read()
    return sorted(set(re.findall(r"\b(fyfft_[a-z_]+)\s*\(", txt)))


def test_exports_every_declared_symbol(fy):
    syms = header_symbols()
    assert len(syms) >= 11
    lib = fy.lib()
    for s in syms:
        assert hasattr(lib, s), s


def test_status_strings(fy):
    lib = fy.lib()
    for code in range(8):
        assert lib.fyfft_status_string(code)
    assert lib.fyfft_status_string(99) == b"unknown status"


def test_argument_checks_without_gpu(fy):
    lib = fy.lib()
    h = ctypes.c_void_p()
    assert lib.fyfft_plan_create(ctypes.byref(h), 0, 1, 0) == 2          # invalid length
    assert lib.fyfft_plan_create(ctypes.byref(h), (1 << 24) + 1, 1, 0) == 2
    assert lib.fyfft_plan_create(ctypes.byref(h), 64, 0, 0) == 1         # batch 0
    assert lib.fyfft_plan_create(ctypes.byref(h), 64, 1, 2) == 3         # bf16: MFMA lengths only
    assert lib.fyfft_plan_create(ctypes.byref(h), 64, 1, 9) == 1         # unknown precision
    assert lib.fyfft_plan_create(None, 64, 1, 0) == 1
    assert lib.fyfft_execute(None, None, None, 0, None) == 6             # invalid plan
    assert lib.fyfft_plan_destroy(None) == 0
    s = ctypes.c_int()
    assert lib.fyfft_length_strategy(0, ctypes.byref(s)) == 2
    assert lib.fyfft_length_strategy(16, None) == 1


@pytest.mark.parametrize("n,expect", [
    (1024, "kernel_direct"), (4096, "kernel_direct"), (78125, "kernel_direct"),
    (6720, "mix_radix"), (12, "mix_radix"), (1 << 20, "mix_radix"), (80000, "mix_radix"),
    (65537, "rader"), (11, "rader"), (1009, "rader"),
    (10007, "bluestein"), (1000003, "bluestein"), (10006, "bluestein"), (39062, "bluestein"),
])
def test_length_strategy(fy, n, expect):
    assert fy.length_strategy(n).name == expect


def test_radices_factor_every_smooth_length(fy):
    for n in list(range(2, 2000)) + [4096, 6720, 16384, 19683, 78125, 1 << 20]:
        smooth = all(True for _ in [0])
        m = n
        for p in (2, 3, 5, 7):
            while m % p == 0:
                m //= p
        if m != 1:
            with pytest.raises(fy.FyfftError):
                fy.length_radices(n)
            continue
        r = fy.length_radices(n)
        assert all(x in (2, 3, 4, 5, 7, 8, 9, 16) for x in r)
        prod = 1
        for x in r:
            prod *= x
        assert prod == n
    assert fy.length_radices(4096) == [16, 16, 16]
    assert fy.length_radices(6720) == [8, 8, 7, 5, 3]
    assert fy.length_radices(729) == [9, 9, 9]        # threes paired as radix 9
    assert fy.length_radices(2187) == [9, 9, 9, 3]


def test_invoker_mirror_strategy(fy):
    S = fy.FFT1D_C2C_invoker.acquired_execution_strategy
    assert S(1024) == fy.FFT_execution_strategy.kernel_direct
    assert S(65537) == fy.FFT_execution_strategy.rader
    with pytest.raises(ValueError):
        fy.FFT1D_C2C_invoker(1)


def test_wisdom_text_roundtrip_and_validation(fy):
    """Wisdom text, version 2: entries follow the device line of the GPU model
    they were tuned on; malformed text is refused and leaves the wisdom as it
    was.  (Without a GPU any device line is accepted; on a GPU a device line
    of another model is refused: test_gpu_forms.)"""
    fy.wisdom_forget()
    assert fy.wisdom_export() == "fyfft-wisdom 2\n"
    text = ("fyfft-wisdom 2\ndevice gfx950 256\n"
            "lds 6720 0 0 16 4 1 8x8x7x5x3\n"
            "ldsg 1000 0 0 16 4 4 8x5x5x5\n"
            "fs 100000 0 0 250 16 16\n"
            "fs 65537 0 1 256 16 8\n"
            "bluestein 10007 0 20480\n")
    fy.wisdom_import(text)
    out = fy.wisdom_export()
    assert out.startswith("fyfft-wisdom 2\ndevice gfx950 256\n"), out
    for line in text.splitlines()[2:]:
        assert line in out.splitlines(), (line, out)
    for bad in ["nonsense\n",
                "fyfft-wisdom 1\nlds 6720 0 0 16 4 1 8x8x7x5x3\n",              # version 1: no device identity
                "fyfft-wisdom 2\nlds 6720 0 0 16 4 1 8x8x7x5x3\n",              # entry before a device line
                "fyfft-wisdom 2\ndevice gfx950 256\nlds 6720 0 0 16 4 1 8x8x7x5\n",   # product != N
                "fyfft-wisdom 2\ndevice gfx950 256\nlds 64 0 0 12 4 1 8x8\n",         # bad points/thread
                "fyfft-wisdom 2\ndevice gfx950 256\nlds 64 0 0 16 4 1 6x8\n",         # radix 6 not a codelet
                "fyfft-wisdom 2\ndevice gfx950 256\nfs 100000 0 0 300 16 16\n",       # 300 does not divide
                "fyfft-wisdom 2\ndevice gfx950 256\nfs 65537 0 1 256 12 8\n",         # tile not a power of two
                "fyfft-wisdom 2\ndevice gfx950 256\nbluestein 10007 0 11000\n",       # M < 2N - 1
                "fyfft-wisdom 2\ndevice gfx950\nlds 64 0 0 16 4 1 8x8\n"]:            # no CU count
        with pytest.raises(fy.FyfftError):
            fy.wisdom_import(bad)
    assert fy.wisdom_export() == out
    fy.wisdom_forget()


def test_options_reserved_must_be_zero(fy):
    import ctypes
    opts = fy._Options(0)
    opts.reserved[3] = 1
    h = ctypes.c_void_p()
    assert fy.lib().fyfft_plan_create_ex(ctypes.byref(h), 64, 1, 0, ctypes.byref(opts)) == 1
    opts = fy._Options(7)
    assert fy.lib().fyfft_plan_create_ex(ctypes.byref(h), 64, 1, 0, ctypes.byref(opts)) == 1


def test_real_plan_argument_checks_without_gpu(fy):
    lib = fy.lib()
    h = ctypes.c_void_p()
    assert lib.fyfft_plan_create_real(ctypes.byref(h), 0, 1, 0) == 2     # invalid length
    assert lib.fyfft_plan_create_real(ctypes.byref(h), 64, 0, 0) == 1    # batch 0
    assert lib.fyfft_plan_create_real(ctypes.byref(h), 64, 1, 2) == 3    # bf16 real: unsupported
    assert lib.fyfft_plan_create_real(ctypes.byref(h), 64, 1, 9) == 1    # unknown precision
    assert lib.fyfft_plan_create_real(None, 64, 1, 0) == 1


def test_strided_plan_argument_checks_without_gpu(fy):
    lib = fy.lib()
    h = ctypes.c_void_p()
    lay = fy._Layout(1, 64, 1, 64)
    assert lib.fyfft_plan_create_many(ctypes.byref(h), 64, 0, ctypes.byref(lay), 0) == 1   # howmany 0
    assert lib.fyfft_plan_create_many(ctypes.byref(h), 0, 4, ctypes.byref(lay), 0) == 2    # length 0
    assert lib.fyfft_plan_create_many(ctypes.byref(h), 64, 4, None, 0) == 1                # no layout
    zero = fy._Layout(0, 64, 1, 64)
    assert lib.fyfft_plan_create_many(ctypes.byref(h), 64, 4, ctypes.byref(zero), 0) == 1  # stride 0
    assert lib.fyfft_plan_create_many(ctypes.byref(h), 64, 4, ctypes.byref(lay), 2) == 3   # bf16
    dims = (ctypes.c_size_t * 3)(4, 4, 4)
    assert lib.fyfft_plan_create_nd(ctypes.byref(h), 0, dims, 1, 0) == 1                   # rank 0
    assert lib.fyfft_plan_create_nd(ctypes.byref(h), 4, dims, 1, 0) == 1                   # rank 4
    assert lib.fyfft_plan_create_nd(ctypes.byref(h), 3, dims, 0, 0) == 1                   # batch 0
    bad = (ctypes.c_size_t * 2)(4, 0)
    assert lib.fyfft_plan_create_nd(ctypes.byref(h), 2, bad, 1, 0) == 2                   # dim 0


def test_mix_modes_argument_checks_without_gpu(fy):
    import ctypes
    L = fy.lib()
    null = ctypes.c_void_p(0)
    assert L.fyfft_mix_modes(0, 0, null, null, null, 0, 4, 4, 4, null) == 0      # empty: nothing to do
    assert L.fyfft_mix_modes(9, 0, null, null, null, 1, 4, 4, 4, null) == 1      # unknown kind
    assert L.fyfft_mix_modes(0, 2, null, null, null, 1, 4, 4, 4, null) == 3      # BF16: unsupported
    assert L.fyfft_mix_modes(0, 1, null, null, null, 1, 4, 4, 4, null) == 1      # FP64, null operands
    assert L.fyfft_mix_modes(0, 0, null, null, null, 1, 4, 4, 4, null) == 1      # null operands
    assert L.fyfft_mix_modes(0, 0, ctypes.c_void_p(8), ctypes.c_void_p(8), ctypes.c_void_p(8),
                             70000, 4, 4, 4, null) == 1                          # k > 65535


def test_strategy_processable_matches_reference_predicates(orc):
    """fyfft_strategy_processable answers the reference's own is_processable
    predicates, restated in the oracle (radix_wrap.hpp:93-98,
    mix_radix.hpp:618-643, rader.hpp:403-423, bluestein.hpp:212-215)."""
    import ctypes
    import fyfft_amd as fy
    L = fy.lib()
    L.fyfft_strategy_processable.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int)]
    ok = ctypes.c_int()

    def q(strategy, n):
        assert L.fyfft_strategy_processable(strategy, n, ctypes.byref(ok)) == 0
        return bool(ok.value)
    for n in list(range(2, 2000)) + [6720, 10006, 10007, 39041, 39061, 39062, 39063, 39064, 39067, 65537, 78125]:
        assert q(0, n) == orc.is_directly_supported(n), n
        assert q(1, n) == orc.mix_is_processable(n), n
        assert q(2, n) == orc.rader_is_processable(n), n
        assert q(3, n) == orc.bluestein_is_processable(n), n
        assert q(4, n)
    assert L.fyfft_strategy_processable(9, 16, ctypes.byref(ok)) == 1
