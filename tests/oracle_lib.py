"""ctypes binding of the oracle (oracle/liboracle.so) -- test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
as the checker; the product never does.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

# prime triples: (p, generator)
REF_PRIMES = ((4293918721, 19), (4253024257, 5), (4231004161, 7))      # transform.h:26-29, :128-130
ENGINE_PRIMES = ((4253024257, 5), (4076863489, 7), (4013948929, 19))   # 3*2^22 | p - 1

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
        L = ctypes.CDLL(LIB_PATH)
        u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
        i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
        u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
        c = ctypes
        L.or_powmod32.argtypes = [c.c_uint32, c.c_uint64, c.c_uint32]; L.or_powmod32.restype = c.c_uint32
        L.or_transform_tables.argtypes = [c.c_uint32, c.c_uint32, c.c_uint32, u32p, u32p]
        L.or_square_mod.argtypes = [c.c_uint32, c.c_uint32, u32p, u32p, u32p]
        L.or_mul_mod.argtypes = [c.c_uint32, c.c_uint32, u32p, u32p, u32p, u32p]
        L.or_poly_sqr_algo.argtypes = [c.c_uint32, u32p]
        L.or_garner3_words.argtypes = [c.c_uint32] * 6 + [c.POINTER(c.c_uint64), c.POINTER(c.c_int64)]
        L.or_engine_new.argtypes = [u32p, u32p, c.c_uint32, c.c_uint32, c.c_uint32]; L.or_engine_new.restype = c.c_void_p
        L.or_engine_free.argtypes = [c.c_void_p]
        L.or_engine_square_dup.argtypes = [c.c_void_p, c.c_int, i32p]
        L.or_engine_mul.argtypes = [c.c_void_p, i32p, i32p]
        L.or_engine_square_coefs.argtypes = [c.c_void_p, i32p, u64p, i64p]
        L.or_get_int.argtypes = [c.c_uint32, c.c_uint32, i32p, u32p]; L.or_get_int.restype = c.c_int
        L.or_is_prime.argtypes = [c.c_uint32, c.c_uint32, u32p, c.POINTER(c.c_int), c.POINTER(c.c_uint64)]
        L.or_is_prime.restype = c.c_int
        L.or_exponent.argtypes = [c.c_uint32, c.c_uint32, u32p]; L.or_exponent.restype = c.c_size_t
        L.or_prp.argtypes = [c.c_void_p, c.c_int, c.c_int, u32p, c.POINTER(c.c_uint64), c.POINTER(c.c_int), c.POINTER(c.c_int)]
        L.or_prp.restype = c.c_int
        L.or_prp_steps.argtypes = [c.c_void_p, c.c_int, c.c_int, i32p]
        L.or_prp_steps_window.argtypes = [c.c_void_p, u32p, c.c_int, c.c_int, c.c_int, i32p]
        L.or_jacobi.argtypes = [c.c_uint64, c.c_uint64]; L.or_jacobi.restype = c.c_int
        L.or_isprime64.argtypes = [c.c_uint64]; L.or_isprime64.restype = c.c_int
        L.or_sieve_window.argtypes = [c.c_int, c.c_uint64, c.c_uint64, c.c_uint64, c.c_uint64,
                                      c.c_uint64, c.c_uint64, c.c_int, u8p]
        L.or_sieve_window.restype = c.c_int
        L.or_sieve_factors.argtypes = [c.c_int, c.c_uint64, c.c_uint64, c.c_int, u64p, c.c_uint64, c.POINTER(c.c_uint64)]
        L.or_sieve_factors.restype = c.c_int
        L.or_sieve_products.argtypes = [c.c_int, c.POINTER(c.c_uint64), c.POINTER(c.c_uint64)]
        L.or_powm.argtypes = [c.c_char_p, c.c_size_t, c.c_char_p, c.c_size_t, c.c_char_p, c.c_size_t,
                              c.c_char_p, c.c_size_t]
        L.or_powm.restype = c.c_int
        _lib = L
    return _lib


def tables(p, g, n):
    w = np.zeros(n, np.uint32); wi = np.zeros(n, np.uint32)
    lib().or_transform_tables(p, g, n, w, wi)
    return w, wi


def garner(r1, r2, r3, p1, p2, p3):
    lo = ctypes.c_uint64(); hi = ctypes.c_int64()
    lib().or_garner3_words(r1, r2, r3, p1, p2, p3, ctypes.byref(lo), ctypes.byref(hi))
    return (hi.value << 64) + lo.value


class Engine:
    """One candidate: N = b^n - b^(n/2) + 1, signed digit state."""

    def __init__(self, n, b, csize=16, primes=ENGINE_PRIMES):
        self.n, self.b, self.csize = n, b, csize
        p = np.array([q for q, _ in primes], np.uint32)
        g = np.array([h for _, h in primes], np.uint32)
        self.h = lib().or_engine_new(p, g, n, csize, b)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_engine_free(self.h)
            self.h = None

    def square_dup(self, digits, dup):
        lib().or_engine_square_dup(self.h, int(dup), digits)

    def mul(self, x, y):
        lib().or_engine_mul(self.h, x, y)

    def square_coefs(self, digits):
        lo = np.zeros(self.n, np.uint64); hi = np.zeros(self.n, np.int64)
        lib().or_engine_square_coefs(self.h, digits, lo, hi)
        return [(int(h) << 64) + int(l) for l, h in zip(lo, hi)]

    def prp(self, i0=0, flip_at=-1):
        canon = np.zeros(self.n, np.uint32)
        r64 = ctypes.c_uint64(); prm = ctypes.c_int(); gl = ctypes.c_int()
        rc = lib().or_prp(self.h, i0, flip_at, canon, ctypes.byref(r64), ctypes.byref(prm), ctypes.byref(gl))
        return rc, canon, r64.value, bool(prm.value), bool(gl.value)

    def prp_steps(self, steps, i0=0):
        d = np.zeros(self.n, np.int32)
        lib().or_prp_steps(self.h, i0, steps, d)
        return d


    def prp_steps_window(self, top_value, i_lo, i0, steps):
        """top_value = E >> i_lo (the window of E's bits from i_lo up)."""
        nbytes = (top_value.bit_length() + 7) // 8 + 4
        win = np.frombuffer(top_value.to_bytes(nbytes + (-nbytes) % 4, "little"), np.uint32).copy()
        d = np.zeros(self.n, np.int32)
        lib().or_prp_steps_window(self.h, win, i_lo, i0, steps, d)
        return d


def top_bits_of_exponent(b, n, k):
    """(E >> s, s) with E = b^n - b^(n/2), n a power of two, E >> s holding the
    top k bits exactly (certified truncated squaring; raises when unsure)."""
    P = k + 96
    m = n.bit_length() - 1
    lo, hi, e = b, b, 0          # b^(2^i) in [lo, hi] * 2^e
    for _ in range(m):
        lo, hi, e = lo * lo, hi * hi, 2 * e
        sh = max(0, hi.bit_length() - P)
        if sh:
            lo, hi, e = lo >> sh, (hi >> sh) + 1, e + sh
    # E = b^n - b^(n/2): b^(n/2) < 2^(bits(b^n)/2 + 1) is far below the window
    top_len = hi.bit_length()
    assert lo.bit_length() == top_len, "uncertain length"
    s_local = top_len - k
    a, c = lo >> s_local, hi >> s_local
    assert a == c, "top bits uncertain"
    # b^n = a 2^s + r with r >= t 2^e (t = lo mod 2^s_local): the subtraction
    # of b^(n/2) < 2^((n/2) bits(b)) cannot borrow into the window when t has
    # a bit at or above 2^((n/2) bits(b) + 1 - e)
    t = lo & ((1 << s_local) - 1)
    assert t != 0 and e + t.bit_length() - 1 >= (n // 2) * b.bit_length() + 1, "window too close to b^(n/2)"
    return a, e + s_local


def get_int(n, b, digits):
    canon = np.zeros(n, np.uint32)
    rc = lib().or_get_int(n, b, np.ascontiguousarray(digits, np.int32), canon)
    return rc, canon


def is_prime(n, b, canon):
    prm = ctypes.c_int(); r64 = ctypes.c_uint64()
    err = lib().or_is_prime(n, b, np.ascontiguousarray(canon, np.uint32), ctypes.byref(prm), ctypes.byref(r64))
    return err, bool(prm.value), r64.value


def exponent(n, b):
    w = np.zeros(n + 2, np.uint32)
    ws = lib().or_exponent(n, b, w)
    return sum(int(w[k]) << (32 * k) for k in range(ws))


def sieve_window(m, b_min, b_max, p_max_pos, p_max_neg, mode=0, p_min_pos=None, p_min_neg=11):
    if p_min_pos is None:
        p_min_pos = 3 * (1 << m) + 1
    struck = np.zeros(b_max - b_min + 1, np.uint8)
    rc = lib().or_sieve_window(m, b_min, b_max, p_min_pos, p_max_pos, p_min_neg, p_max_neg, mode, struck)
    assert rc == 0
    return struck


def sieve_products(reset=True):
    """(Montgomery products, verification products) counted since the last reset."""
    a, v = ctypes.c_uint64(), ctypes.c_uint64()
    lib().or_sieve_products(1 if reset else 0, ctypes.byref(a), ctypes.byref(v))
    return a.value, v.value


def sieve_factors(m, p_min, p_max, mode, cap=1 << 20):
    """Oracle factor emission: sorted array of (p, b) rows."""
    pairs = np.zeros(2 * cap, np.uint64)
    cnt = ctypes.c_uint64()
    rc = lib().or_sieve_factors(m, p_min, p_max, mode, pairs, cap, ctypes.byref(cnt))
    assert rc == 0 and cnt.value <= cap
    out = pairs[:2 * cnt.value].reshape(-1, 2)
    return out[np.lexsort((out[:, 1], out[:, 0]))]


# ---- pure-Python helpers (independent of the oracle) ----

def powm(a, e, m):
    """a^e mod m by GMP's mpz_powm (oracle/gmp_powm.c)."""
    nb = max(1, (a.bit_length() + 7) // 8)
    ne = max(1, (e.bit_length() + 7) // 8)
    nm = max(1, (m.bit_length() + 7) // 8)
    out = ctypes.create_string_buffer(nm)
    rc = lib().or_powm(a.to_bytes(nb, "little"), nb, e.to_bytes(ne, "little"), ne, m.to_bytes(nm, "little"), nm,
                       out, nm)
    assert rc == 0
    return int.from_bytes(out.raw, "little")


def digits_value(digits, b):
    """sum d_j b^j (divide and conquer: quasi-linear with Python's big integers)."""
    ds = [int(x) for x in digits]
    powers = {}

    def bpow(k):
        if k not in powers:
            powers[k] = pow(b, k)
        return powers[k]

    def rec(lo, hi):
        if hi - lo <= 64:
            v = 0
            for d in reversed(ds[lo:hi]):
                v = v * b + d
            return v
        mid = (lo + hi) // 2
        return rec(lo, mid) + rec(mid, hi) * bpow(mid - lo)

    return rec(0, len(ds)) if ds else 0


def modulus(n, b):
    return b ** n - b ** (n // 2) + 1


def canon_digits(v, n, b):
    out = []
    for _ in range(n):
        v, r = divmod(v, b)
        out.append(r)
    assert v == 0
    return out
