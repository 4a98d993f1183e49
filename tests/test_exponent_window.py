"""Top bits of the PRP exponent E = N - 1 (cyclo.h:258-271) for both twin
members, from the certified truncated squaring (ctwin_amd.exponent.top_bits
and the oracle's top_bits_of_exponent), against E computed exactly."""
import random

import pytest

import oracle_lib as O
from ctwin_amd.exponent import exponent, top_bits, window_value


def exact_top(b, n, k, sign):
    E = b ** n - b ** (n // 2) - (0 if sign > 0 else 2)
    s = E.bit_length() - k
    return E >> s, s


@pytest.mark.parametrize("sign", [1, -1])
@pytest.mark.parametrize("n", [64, 256, 1024, 4096])
def test_top_bits_equal_exact(n, sign):
    rng = random.Random(n * 7 + sign)
    for _ in range(12):
        b = rng.randrange(33, 1 << 31)
        h = (n // 2) * b.bit_length()
        for k in (16, 64, 200, 1000):
            if k + 200 > (n * b.bit_length() - h) // 2:
                continue
            want = exact_top(b, n, k, sign)
            assert top_bits(b, n, k, sign) == want, (b, n, k, sign)
            if sign > 0:
                assert O.top_bits_of_exponent(b, n, k) == want, (b, n, k)
            v, top = window_value(b, n, k, sign)
            assert top == exponent(b, n, sign).bit_length() - 1
            assert v >> (top + 1 - k) == want[0]


def test_twin_members_share_the_window():
    # E- = E+ - 2 differs from E+ only below the window
    rng = random.Random(3)
    for n in (1024, 4096):
        for _ in range(8):
            b = rng.randrange(33, 1 << 20)
            assert top_bits(b, n, 300, 1) == top_bits(b, n, 300, -1)


def test_window_near_the_subtracted_half_is_refused():
    # n = 64, b = 33: E has 323 bits and b^32 has 162, so a 200-bit window
    # ends below bit 192 = 32 bits(b), where the subtraction may borrow
    with pytest.raises(ArithmeticError):
        top_bits(33, 64, 200)
    with pytest.raises(AssertionError):
        O.top_bits_of_exponent(33, 64, 200)
