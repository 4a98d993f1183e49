"""Exponent bits of E = N - 1 = b^n - b^(n/2) for the PRP loop (cyclo.h:258-271).

The full exponent has n log2(b) bits (35 Mbit at n = 2^21); a run of t
squarings only needs its top t bits.  `top_bits` returns them exactly from a
certified truncated squaring of b (interval arithmetic on the mantissa), and
raises if the interval cannot certify them.
"""


def exponent(b, n):
    h = b ** (n // 2)
    return h * h - h


def top_bits(b, n, k):
    """Returns (w, s) with w = E >> s holding exactly the top k bits of E."""
    if n < 64:
        e = exponent(b, n)
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
    # b^(n/2) lies far below the window; the subtraction can only reach it
    # through a borrow chain across all-zero bits below the window
    if e + s_local <= (n // 2) * b.bit_length() + 64 or (lo & ((1 << s_local) - 1)) == 0:
        raise ArithmeticError("exponent window too close to b^(n/2)")
    return w, e + s_local


def window_value(b, n, k):
    """E truncated to its top k bits, as an int of E's bit length (lower bits 0)."""
    w, s = top_bits(b, n, k)
    return w << s, s + k - 1
