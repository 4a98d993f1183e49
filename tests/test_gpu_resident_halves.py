"""Resident kernels with the two primes of a squaring in the halves of one
workgroup (transform lengths 2^10 .. 2^13): the same carried digits as the
one-prime-after-the-other kernel over PRP-loop runs with their exponent bits
(N+ and N-, bases 2 and 3, several lanes), and the oracle's digits at
n = 2^10; a Gerbicz-Li product between runs keeps both forms equal."""
import random

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
ct = pytest.importorskip("ctwin_amd")


@pytest.mark.parametrize("ln,sign,base,lanes", [
    (10, 1, 2, 5), (11, 1, 3, 3), (12, 1, 2, 4), (13, 1, 2, 2),
    (9, -1, 3, 3), (12, -1, 2, 2),
])
def test_halves_equal_sequential(ln, sign, base, lanes):
    n = 1 << ln
    rng = random.Random(4200 + ln + 7 * lanes)
    bs = [2 * rng.randrange(1 << 14, 1 << 16) for _ in range(lanes)]
    E = [b ** n - b ** (n // 2) - (2 if sign < 0 else 0) for b in bs]     # N - 1
    i0 = max(e.bit_length() for e in E) - 1
    print(f"\n[case] halves ln={ln} sign={sign} base={base} lanes={lanes} bs={bs}", flush=True)
    engines = []
    for halves in (True, False):
        t = ct.Transform(ln, lanes, ct.KIND_RESIDENT, sign=sign)
        t.set_base(base)
        t.set_prime_halves(halves)
        t.set_b(bs); t.set_exponents(E); t.set(1); t.copy(1, 0)
        assert t.primes == 2
        engines.append(t)
    assert engines[0].tuning()[1] and not engines[1].tuning()[1]
    i = i0
    for rep, steps in enumerate((50, 37, 90)):
        for t in engines:
            t.run(i, i - steps + 1)
        i -= steps
        a, b = engines[0].read_state(0), engines[1].read_state(0)
        assert np.array_equal(a, b), f"run {rep}"
        for t in engines:
            t.mul(1, 1, 0)                      # a Gerbicz-Li product (the one-prime-after-the-other kernel)
        assert np.array_equal(engines[0].read_state(1), engines[1].read_state(1)), f"product {rep}"
    if ln == 10 and sign > 0 and base == 2:      # the oracle's base-2 PRP loop
        for l, b in enumerate(bs):
            assert np.array_equal(engines[0].read_state(0)[l], O.Engine(n, b, 16).prp_steps(i0 - i, i0))
    for t in engines:
        t.close()
