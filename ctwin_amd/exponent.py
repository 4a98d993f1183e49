"""Exponent bits of E = N - 1 for the PRP loop (cyclo.h:258-271):
E = b^n - b^(n/2) for N+ = b^n - b^(n/2) + 1, and b^n - b^(n/2) - 2 for
N- = b^n - b^(n/2) - 1.

The full exponent has n log2(b) bits (35 Mbit at n = 2^21); a run of t
squarings only needs its top t bits.  `top_bits` returns them exactly from a
certified truncated squaring of b (interval arithmetic on the mantissa), and
raises if the interval cannot certify them.
"""


def exponent(b, n, sign=1):
    h = b ** (n // 2)
    return h * h - h - (0 if sign > 0 else 2)


def top_bits(b, n, k, sign=1):
    """Returns (w, s) with w = E >> s holding exactly the top k bits of E."""
    if n < 64:
        e = exponent(b, n, sign)
        s = max(0, e.bit_length() - k)
        return e >> s, s
    P = k + 96
    m = n.bit_length() - 1
    lo, hi, e = b, b, 0              # b^(2^i) lies in [lo, hi] * 2^e
    for _ in range(m):
        lo, hi, e = lo * lo, hi * hi, 2 * e
        sh = max(0, hi.bit_length() - P)
        if sh:
            lo, hi, e = lo >> sh, (hi >> sh) + 1, e + sh
    if lo.bit_length() != hi.bit_length():
        raise ArithmeticError("top bits of b^n not certified")
    s_local = hi.bit_length() - k
    w = lo >> s_local
    if w != hi >> s_local:
        raise ArithmeticError("top bits of b^n not certified")
    # b^n = w 2^s + r with r >= t 2^e, t = lo mod 2^s_local; subtracting
    # b^(n/2) (+ 2) < 2^(h + 1), h = (n/2) bits(b), leaves the window alone
    # when r >= 2^(h + 1), i.e. when t has a bit at or above 2^(h + 1 - e)
    t = lo & ((1 << s_local) - 1)
    if t == 0 or e + t.bit_length() - 1 < (n // 2) * b.bit_length() + 1:
        raise ArithmeticError("exponent window too close to b^(n/2)")
    return w, e + s_local


def window_value(b, n, k, sign=1):
    """E truncated to its top k bits, as an int of E's bit length (lower bits 0)."""
    w, s = top_bits(b, n, k, sign)
    return w << s, s + k - 1
