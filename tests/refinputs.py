"""Inputs of the reference's accuracy measurements, regenerated in numpy.

The reference's example program (example.cpp:16-23) fills a transform with
``std::mt19937`` and ``std::uniform_real_distribution<>(-1.0, 1.0)`` (a
double distribution), drawing the real part and then the imaginary part of
each element and casting both to float.  The FP32 errors recorded for the
compiled reference in golden/reference_facts.json were measured on that
stream with seed 0, one transform per length.  libstdc++ draws a double
uniform from two 32-bit outputs (generate_canonical<double, 53>:
(u0 + u1 * 2^32) / 2^64, clamped below 1), then scales it to [-1, 1); numpy's
legacy MT19937 seeding of an integer equals std::mt19937's.
"""
import numpy as np


def reference_inputs(n: int, seed: int = 0) -> np.ndarray:
    """complex64 vector of length n drawn as the reference's example does."""
    u = np.random.RandomState(seed)._bit_generator.random_raw(4 * n).astype(np.uint64)
    s = u[0::2].astype(np.float64) + u[1::2].astype(np.float64) * 2.0 ** 32
    r = np.minimum(s / 2.0 ** 64, np.nextafter(1.0, 0.0))
    v = (r * 2.0 - 1.0).astype(np.float32)
    return (v[0::2] + 1j * v[1::2]).astype(np.complex64)
