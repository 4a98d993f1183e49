"""Pins the oracle's squaring-engine restatement (path A) against the
reference's known-answer strings, exact big-integer arithmetic and closed
forms.  CPU only."""
import json
import os
import random
import re

import numpy as np
import pytest

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
ALL_PRIMES = sorted(set(O.REF_PRIMES) | set(O.ENGINE_PRIMES))


def parse_pari(s, n):
    """' = 284*x^7 + ... - 291' -> coefficient list c[0..n-1]."""
    s = s.replace(" = ", "").replace(" ", "")
    c = [0] * n
    for sign, mag, xp in re.findall(r"([+-]?)(\d+)(\*x(?:\^\d+)?)?", s):
        e = 0 if not xp else (1 if xp == "*x" else int(xp.split("^")[1]))
        c[e] = int(mag) * (-1 if sign == "-" else 1)
    return c


def py_poly_mul_mod(a, b, n):
    """Independent: big-int schoolbook product, X^n = X^(n/2) - 1."""
    t = [0] * (2 * n)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            t[i + j] += x * y
    for d in range(2 * n - 1, n - 1, -1):
        v = t[d]; t[d] = 0
        t[d - n // 2] += v; t[d - n] -= v
    return t[:n]


@pytest.mark.parametrize("p,g", ALL_PRIMES)
@pytest.mark.parametrize("n", [4, 8, 64, 1024])
def test_tables_identities(p, g, n):
    w, wi = O.tables(p, g, n)
    w = [int(x) for x in w]; wi = [int(x) for x in wi]
    r = w[1]
    assert (r * r - r + 1) % p == 0                       # primitive 6th root
    if n >= 4:   # level-1 factors are X^(n/2) - r and X^(n/2) - r', r' = 1 - r
        assert pow(w[2], 2, p) == r and pow(w[3], 2, p) == (1 - r) % p
    s = 2
    while 2 * s < n:
        for j in range(s):
            assert pow(w[2 * s + 2 * j], 2, p) == w[s + j] % p        # child roots
            assert pow(w[2 * s + 2 * j + 1], 2, p) == (-w[s + j]) % p
        s *= 2
    s = 2
    while s < n:
        for j in range(s):
            assert (w[s + j] * wi[s + j]) % p == 1                    # inverse table
        s *= 2
    assert wi[0] == w[1]
    assert wi[1] == (2 * pow((1 - 2 * r) % p, p - 2, p) * 2 * pow(n, p - 2, p)) % p


def test_algo_known_answer():
    kat = json.load(open(os.path.join(HERE, "golden", "algo_kat.json")))
    p1 = 4293918721
    for case in kat["cases"]:
        n = case["n"]
        want = parse_pari(case["pari"], n)
        # cross-check the fixture itself against exact arithmetic
        assert py_poly_mul_mod(list(range(1, n + 1)), list(range(1, n + 1)), n) == want
        for e in range(1, 5):
            f = 10 ** e
            x = np.array([(i + 1) * f for i in range(n)], np.uint32)
            O.lib().or_poly_sqr_algo(n, x)
            for i in range(n):
                assert int(x[i]) == (want[i] * f * f) % p1
                if max(abs(c) for c in want) * f * f < p1 // 2:    # algo.cpp:162 divide-back
                    v = int(x[i]); v = v - p1 if v > p1 // 2 else v
                    assert v // (f * f) == want[i] and v % (f * f) == 0


@pytest.mark.parametrize("p,g", ALL_PRIMES)
@pytest.mark.parametrize("n", [4, 8, 16, 256, 1024])
def test_square_and_mul_match_exact(p, g, n):
    rng = random.Random(n * 7 + p % 1000)
    w, wi = O.tables(p, g, n)
    for _ in range(3):
        a = [rng.randrange(-(1 << 20), 1 << 20) for _ in range(n)]
        c = [rng.randrange(-(1 << 20), 1 << 20) for _ in range(n)]
        want_sq = [v % p for v in py_poly_mul_mod(a, a, n)]
        x = np.array([v % p for v in a], np.uint32)
        O.lib().or_square_mod(p, n, w, wi, x)
        assert [int(v) for v in x] == want_sq
        want_mul = [v % p for v in py_poly_mul_mod(a, c, n)]
        x = np.array([v % p for v in a], np.uint32)
        y = np.array([v % p for v in c], np.uint32)
        y0 = y.copy()
        O.lib().or_mul_mod(p, n, w, wi, x, y)
        assert [int(v) for v in x] == want_mul
        assert np.array_equal(y, y0)


def test_garner_matches_python_crt():
    rng = random.Random(5)
    for primes in (O.REF_PRIMES, O.ENGINE_PRIMES):
        p1, p2, p3 = [q for q, _ in primes]
        M = p1 * p2 * p3
        for _ in range(2000):
            v = rng.randrange(-(M // 2) + 1, M // 2 + 1)
            got = O.garner(v % p1, v % p2, v % p3, p1, p2, p3)
            assert got == v
        # boundaries of (-M/2, M/2]
        for v in (M // 2, -(M // 2), 0, 1, -1):
            assert O.garner(v % p1, v % p2, v % p3, p1, p2, p3) == v


@pytest.mark.parametrize("n,b,csize", [(16, 40, 2), (16, 1000003, 8), (64, (1 << 20) + 5, 16),
                                       (1024, (1 << 19) + 38, 16), (1024, 2000000, 64), (256, 65538, 8)])
def test_engine_square_dup_value(n, b, csize):
    rng = random.Random(n + b)
    N = O.modulus(n, b)
    e_ref = O.Engine(n, b, csize, O.REF_PRIMES)
    e_eng = O.Engine(n, b, csize, O.ENGINE_PRIMES)
    for dup in (0, 1, 1, 0):
        d = np.array([rng.randrange(-b + 1, b) for _ in range(n)], np.int32)
        v = O.digits_value(d, b)
        d1 = d.copy(); d2 = d.copy()
        e_ref.square_dup(d1, dup); e_eng.square_dup(d2, dup)
        assert np.array_equal(d1, d2)                       # prime-set independence
        assert O.digits_value(d1, b) % N == (v * v * (2 if dup else 1)) % N
        # carried digits: |d| < b except the chunk-last digit (normalize2 :1077-1079)
        inner = np.array([k for k in range(n) if k % csize != csize - 1])
        assert np.all(np.abs(d1[inner].astype(np.int64)) < b)
        # exact coefficients before the carry equal the schoolbook product
        coefs = e_eng.square_coefs(d)
        assert coefs == py_poly_mul_mod([int(x) for x in d], [int(x) for x in d], n)


@pytest.mark.parametrize("n,b,csize", [(16, 977, 4), (1024, (1 << 19) + 1234, 16)])
def test_engine_mul_value(n, b, csize):
    rng = random.Random(b)
    N = O.modulus(n, b)
    e = O.Engine(n, b, csize)
    x = np.array([rng.randrange(-b + 1, b) for _ in range(n)], np.int32)
    y = np.array([rng.randrange(-b + 1, b) for _ in range(n)], np.int32)
    vx, vy = O.digits_value(x, b), O.digits_value(y, b)
    y0 = y.copy()
    e.mul(x, y)
    assert O.digits_value(x, b) % N == (vx * vy) % N
    assert np.array_equal(y, y0)


def test_normalize_tiny_hand_case():
    # n = 4, csize = 2, b = 10, all coefficients 0 except one -> exact carries
    e = O.Engine(4, 10, 2)
    d = np.array([7, 0, 0, 0], np.int32)          # 7^2 = 49 -> digits 9, 4
    e.square_dup(d, 0)
    assert list(d) == [9, 4, 0, 0]
    d = np.array([7, 0, 0, 0], np.int32)
    e.square_dup(d, 1)                            # 98 -> 8, 9
    assert list(d) == [8, 9, 0, 0]
    d = np.array([0, 0, 0, 9], np.int32)          # (9 X^3)^2 = 81 X^6 = 81 (X^4 - X^2) ... mod X^4 - X^2 + 1
    e.square_dup(d, 0)                            # X^6 = -1 -> -81: chunk0 gets -81 -> digits -1, -8
    assert O.digits_value(d, 10) % O.modulus(4, 10) == (9 * 1000) ** 2 % O.modulus(4, 10)


@pytest.mark.parametrize("n,b", [(16, 1000), (64, (1 << 20) + 2), (1024, (1 << 19) + 100)])
def test_get_int_and_res64(n, b):
    rng = random.Random(n)
    N = O.modulus(n, b)
    cases = [rng.randrange(-b + 1, b) for _ in range(n)]
    vecs = [np.array(cases, np.int32)]
    vecs.append(np.array(O.canon_digits(N - 1, n, b), np.int32))          # N - 1
    vecs.append(np.array(O.canon_digits(N, n, b), np.int32))              # N -> 0
    vecs.append(np.array(O.canon_digits(N + 5, n, b), np.int32))          # overflow branch
    top = [b - 1] * (n // 2); low = [0] * (n // 2); low[3] = 2
    vecs.append(np.array(low + top, np.int32))                             # top half all b - 1
    vecs.append(-np.array(O.canon_digits(12345, n, b), np.int32))          # negative value
    for d in vecs:
        rc, canon = O.get_int(n, b, d)
        assert rc == 0
        want = O.digits_value(d, b) % N
        assert O.digits_value(canon, b) == want
        assert all(0 <= int(c) < b for c in canon)
        err, prm, r64 = O.is_prime(n, b, canon)
        assert r64 == want % (1 << 64)
        assert prm == (want == 1) and err == (want == 0)


@pytest.mark.parametrize("n,b", [(16, 33), (64, 100003), (1024, 786434)])
def test_exponent(n, b):
    assert O.exponent(n, b) == b ** n - b ** (n // 2)


def find_prp(n, lo, count):
    from sympy import isprime
    out = []
    b = lo
    while len(out) < count:
        if isprime(O.modulus(n, b)):
            out.append(b)
        b += 2
    return out


@pytest.mark.parametrize("n,csize", [(16, 4), (32, 8), (64, 16)])
def test_prp_matches_pow(n, csize):
    rng = random.Random(n)
    bs = [rng.randrange(1 << 15, 1 << 16) * 2 for _ in range(3)] + find_prp(n, 1000, 1)
    for b in bs:
        N = O.modulus(n, b)
        rc, canon, r64, prm, gl = O.Engine(n, b, csize).prp()
        want = pow(2, N - 1, N)
        assert rc in (0,) and gl
        assert O.digits_value(canon, b) == want
        assert r64 == want % (1 << 64)
        assert prm == (want == 1)


def test_prp_n2e10_size():
    """One base-2 PRP at n = 1024 digits (b even in [2^19, 2^20))."""
    rng = random.Random(1)
    b = rng.randrange(1 << 18, 1 << 19) * 2
    n = 1024
    N = O.modulus(n, b)
    rc, canon, r64, prm, gl = O.Engine(n, b, 16).prp()
    assert rc == 0 and gl
    assert O.digits_value(canon, b) == pow(2, N - 1, N)


def test_prp_gerbicz_li_detects_fault_and_i0_independence():
    n, b = 32, 70002
    N = O.modulus(n, b)
    rc, canon, r64, prm, gl = O.Engine(n, b, 8).prp(flip_at=200)
    assert not gl
    assert O.digits_value(canon, b) != pow(2, N - 1, N)
    top = (b ** n - b ** (n // 2)).bit_length() - 1
    rc, canon, r64, prm, gl = O.Engine(n, b, 8).prp(i0=top + 37)
    assert gl and O.digits_value(canon, b) == pow(2, N - 1, N)


def test_prp_steps_are_powers_of_two():
    n, b = 64, 123456
    N = O.modulus(n, b)
    E = b ** n - b ** (n // 2)
    top = E.bit_length() - 1
    for t in (1, 5, 100, 1000):
        d = O.Engine(n, b, 16).prp_steps(t)
        want = pow(2, E >> (top - t + 1), N)
        assert O.digits_value(d, b) % N == want


def test_gmp_powm_oracle_matches_python_pow():
    """The GMP mpz_powm oracle used by the GPU tests for large PRP residues
    equals Python's pow on random and structured cases."""
    import random
    rng = random.Random(5)
    for _ in range(50):
        m = rng.randrange(2, 1 << rng.randrange(2, 700)) | 1
        a, e = rng.randrange(0, 1 << 300), rng.randrange(0, 1 << 600)
        assert O.powm(a, e, m) == pow(a, e, m)
    for sign in (1, -1):
        for b in (1000, 65537, 700001):
            N = b ** 64 - b ** 32 + sign
            assert O.powm(3, N - 1, N) == pow(3, N - 1, N)
    assert O.powm(5, 0, 7) == 1 and O.powm(0, 5, 7) == 0
