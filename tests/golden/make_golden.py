"""Writes the golden fixtures under tests/golden/.

* config1_n1024.npz -- the N=1024 config: x with re/im i.i.d. uniform[-1,1)
  from numpy.random.default_rng(0) (complex64), and its FP64 naive-DFT forward
  transform and inverse, from the oracle's long-double naive DFT.
* reference_facts.json is NOT written here: it holds planner integers that
  were measured on the compiled reference (strategy per N, six-step
  decompositions, Rader/Bluestein pad lengths, FP32 error levels) and is
  committed as data.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def config1():
    n = 1024
    rng = np.random.default_rng(0)
    re = rng.uniform(-1.0, 1.0, n)
    im = rng.uniform(-1.0, 1.0, n)
    x = (re + 1j * im).astype(np.complex64)
    fwd = oracle.dft_naive(x.astype(np.complex128), inverse=False)
    inv = oracle.dft_naive(x.astype(np.complex128), inverse=True)
    np.savez(os.path.join(HERE, "config1_n1024.npz"), x=x, fwd=fwd, inv=inv)


if __name__ == "__main__":
    config1()
    print("wrote", os.path.join(HERE, "config1_n1024.npz"))
