"""fyfft_execute_noalloc (the entry point that allocates nothing) against
fyfft_execute on the same plans, through the C ABI.

Both run the same kernels on the same inputs; only where the four-step
workspace lives (the library's static scratch array instead of the plan's
allocation) and how many transforms a chunk holds differ, neither of which
changes any transform's arithmetic, so the outputs must be bit-identical.
Consecutive calls reuse the scratch array with different plans, the later
ones smaller, so stale contents from an earlier call would show."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(fy, plan, x, y, direction, stream):
    return fy.lib().fyfft_execute_noalloc(plan._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                                          int(direction), ctypes.c_void_p(stream))


CALLS = [
    # (n, batch, precision, direction, describe prefix): larger first, then smaller
    (1 << 20, 20, "f32", 0, "fourstep"),        # 8 MiB per transform: 8 per scratch chunk, 3 chunks
    (65537, 3, "f32", 0, "rader-fourstep"),     # bucketed Rader: tmp + zw + x[0] slots
    (1000003, 2, "f32", 1, "bluestein-fourstep"),
    (30241, 2, "f64", 0, "rader-fourstep"),
    (4096, 7, "f32", 0, "lds-fixed"),           # single-kernel plan: no workspace
]


def test_noalloc_matches_execute(fy, cuda):
    import torch
    stream = torch.cuda.current_stream()
    h = stream.cuda_stream
    for n, batch, prec, direction, kind in CALLS:
        print(f"[noalloc] n={n} batch={batch} {prec} dir={direction}", flush=True)
        rdt = torch.float32 if prec == "f32" else torch.float64
        g = torch.Generator(device="cuda").manual_seed(n + batch)
        x = torch.view_as_complex(torch.randn(batch, n, 2, dtype=rdt, device="cuda", generator=g))
        with fy.Plan(n, batch, fy.F32 if prec == "f32" else fy.F64) as p:
            assert p.describe().startswith(kind), p.describe()
            y_ref = p.execute(x, None, direction)
            # fill the output with garbage so unwritten elements would show
            y = torch.full_like(x, complex(float("nan"), float("nan")))
            st = _noalloc(fy, p, x, y, direction, h)
            assert st == 0, (st, fy.lib().fyfft_last_error())
            torch.cuda.synchronize()
            assert torch.equal(y, y_ref), (n, float((y - y_ref).abs().max()))
            # a second call on the same plan, in place
            z = x.clone()
            assert _noalloc(fy, p, z, z, direction, h) == 0
            torch.cuda.synchronize()
            assert torch.equal(z, y_ref), n


def test_noalloc_limits(fy, cuda):
    import torch
    # one FP32 transform of 2^24 needs 128 MiB of workspace: beyond the 64 MiB scratch
    with fy.Plan(1 << 24, 1, fy.F32) as p:
        x = torch.zeros(1, 1 << 24, dtype=torch.complex64, device="cuda")
        assert _noalloc(fy, p, x, x, 0, 0) == 2  # FYFFT_INVALID_LENGTH
        assert b"static scratch" in fy.lib().fyfft_last_error()
    gp = fy.GroupedPlan(np.array([16, 100]), fy.F32)
    x = torch.zeros(116, dtype=torch.complex64, device="cuda")
    assert fy.lib().fyfft_execute_noalloc(gp._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                                          0, None) == 3  # FYFFT_UNSUPPORTED
    gp.close()
