"""The C++ drop-in (include/foye_FFT.hpp) used the way the reference's
example uses fy::fft::FFT1D_C2C_invoker, compiled with plain g++ against
libfyfft.so (tests/cpp/dropin_main.cpp, built by build())."""
import math
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "fyfft_amd", "lib", "dropin_main")


# invokers: how many of the reference's per-strategy invokers accept N
# (is_processable, the reference's predicates): mixed-radix for 7-smooth N,
# Rader for primes up to 39064 outside the direct table, Bluestein up to
# 39062, DFT always
@pytest.mark.parametrize("n,strategy,invokers", [(1024, 0, 3), (4096, 0, 3), (6720, 1, 3), (17, 2, 3),
                                                 (65537, 2, 1), (10007, 3, 3), (22, 3, 2)])
def test_dropin_header(orc, tmp_path, n, strategy, invokers):
    if n == 1024:
        x = np.load(os.path.join(ROOT, "tests", "golden", "config1_n1024.npz"))["x"]
    else:
        rng = np.random.default_rng(n)
        x = (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(np.complex64)
    fin, ffwd, finv = tmp_path / "in.bin", tmp_path / "fwd.bin", tmp_path / "inv.bin"
    x.view(np.float32).tofile(fin)
    r = subprocess.run([EXE, str(n), str(fin), str(ffwd), str(finv)], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert f"strategy {strategy}" in r.stdout
    assert f"invokers {invokers}" in r.stdout, r.stdout
    fwd = np.fromfile(ffwd, np.float32).view(np.complex64)
    inv = np.fromfile(finv, np.float32).view(np.complex64)
    x64 = x.astype(np.complex128)
    tol = 1e-7 * math.log2(n) + 1e-7
    assert orc.rel_l2(fwd, orc.fft_ext(x64)) < tol
    assert orc.rel_l2(inv, orc.fft_ext(x64, inverse=True)) < tol
