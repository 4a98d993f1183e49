"""The C++ CPU comparator (oracle/cpu_port/ct_cpu.cpp) reproduces the
oracle's digits (the reference's normalize1/normalize2 device digits for
csize 16) step for step: several lanes (parallel over lanes), one lane with
the whole-array levels and in-cache blocks (n > 2^14, parallel inside the
transform), one and four threads alike (CPU only)."""
import random

import numpy as np
import pytest

import cpu_port as C
import oracle_lib as O
from ctwin_amd.exponent import exponent


def _oracle_prp(n, b, x0, E, i0, steps):
    e = O.Engine(n, b, 16)
    x = x0.copy()
    for t in range(steps):
        i = i0 - t
        e.square_dup(x, (E >> i) & 1)
    return x


@pytest.mark.parametrize("ln,lanes,threads", [(6, 3, 1), (10, 4, 4), (14, 2, 2), (15, 1, 4), (16, 1, 1)])
def test_cpu_port_matches_oracle(ln, lanes, threads):
    n = 1 << ln
    rng = random.Random(ln * 7 + lanes)
    bs = [rng.randrange(1 << 15, 1 << 18) for _ in range(lanes)]
    E = [exponent(b, n) for b in bs]
    top = max(e.bit_length() for e in E)
    i0 = top - 1
    x = np.array([[rng.randrange(-b + 1, b) for _ in range(n)] for b in bs], np.int32)
    steps = 3
    want = np.stack([_oracle_prp(n, bs[l], x[l], E[l], i0, steps) for l in range(lanes)])
    bits, words = C.pack_bits(E, lanes)
    eng = C.CpuEngine(ln, bs)
    got = x.copy()
    eng.steps(got, bits, words, i0, steps, threads)
    assert np.array_equal(got, want), f"first mismatch at {np.argwhere(got != want)[:4]}"
    if threads > 1:
        again = x.copy()
        eng.steps(again, bits, words, i0, steps, 1)
        assert np.array_equal(again, got)
