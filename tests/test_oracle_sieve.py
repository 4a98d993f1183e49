"""Pins the oracle's window sieve (path B) against brute force and against the
survivor lists written by the reference program (tests/golden/sieve_small.json).
CPU only."""
import json
import os

import numpy as np
import pytest

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))


def cases():
    return json.load(open(os.path.join(HERE, "golden", "sieve_small.json")))["cases"]


def survivors_of(case):
    d = case["deltas"]
    return list(np.cumsum(np.array(d, dtype=np.int64))) if d else []


def primes_upto(n):
    s = np.ones(n + 1, bool); s[:2] = False
    for i in range(2, int(n ** 0.5) + 1):
        if s[i]:
            s[i * i::i] = False
    return np.nonzero(s)[0]


def k_max_pos(m, p_max):
    k = (p_max >> m) // 3
    while 3 * (k << m) + 1 < p_max:
        k += 1
    return k


def brute_force(m, b_min, b_max, p_max_pos, p_max_neg):
    """Strike b when a prime p of the processed sets divides N+(b) or N-(b):
    y = b^(2^(m-1)) mod p, y^2 - y + 1 = 0 (+) or y^2 - y - 1 = 0 (-)."""
    b = np.arange(b_min, b_max + 1, dtype=np.int64)
    struck = np.zeros(len(b), bool)
    kp = k_max_pos(m, p_max_pos)
    pmax_p = 3 * (kp << m) + 1
    kn = p_max_neg // 10
    while 10 * kn + 1 < p_max_neg:
        kn += 1
    pmax_n = 10 * kn + 1
    for p in primes_upto(max(pmax_p, pmax_n)):
        p = int(p)
        pos = (p - 1) % (3 << m) == 0 and p <= pmax_p and p >= 3 * (1 << m) + 1
        neg = p % 10 in (1, 9) and 11 <= p <= pmax_n
        if not (pos or neg):
            continue
        y = b % p
        for _ in range(m - 1):
            y = (y * y) % p
        if pos:
            struck |= ((y * y - y + 1) % p) == 0
        if neg:
            struck |= ((y * y - y - 1) % p) == 0
    return [int(x) for x in b[~struck]]


def test_oracle_matches_brute_force_tiny():
    c = cases()[0]
    assert c["name"] == "m8_tiny"
    want = brute_force(c["m"], c["b_min"], c["b_max"], c["p_max_pos"], c["p_max_neg"])
    assert want == [int(x) for x in survivors_of(c)]          # the reference agrees with brute force
    st = O.sieve_window(c["m"], c["b_min"], c["b_max"], c["p_max_pos"], c["p_max_neg"])
    got = [c["b_min"] + int(i) for i in np.nonzero(st == 0)[0]]
    assert got == want


@pytest.mark.parametrize("idx", [0, 1, 2])
def test_oracle_matches_reference_survivors(idx):
    c = cases()[idx]
    st = O.sieve_window(c["m"], c["b_min"], c["b_max"], c["p_max_pos"], c["p_max_neg"])
    got = [c["b_min"] + int(i) for i in np.nonzero(st == 0)[0]]
    assert len(got) == c["count"]
    assert got == [int(x) for x in survivors_of(c)]


def test_oracle_sides_separately():
    c = cases()[0]
    pos = O.sieve_window(c["m"], c["b_min"], c["b_max"], c["p_max_pos"], c["p_max_neg"], mode=1)
    neg = O.sieve_window(c["m"], c["b_min"], c["b_max"], c["p_max_pos"], c["p_max_neg"], mode=-1)
    both = O.sieve_window(c["m"], c["b_min"], c["b_max"], c["p_max_pos"], c["p_max_neg"], mode=0)
    assert np.array_equal(both, pos | neg)


# ---------------- factor emission (sieve/GPU semantics) ----------------

def brute_roots(p, m, plus):
    """All b in [0, p) with p | N+(b) (plus) or N-(b), by direct evaluation."""
    b = np.arange(p, dtype=np.int64)
    y = b.copy()
    for _ in range(m - 1):
        y = (y * y) % p
    v = (y * y - y + (1 if plus else -1)) % p
    return set(int(x) for x in b[v == 0])


@pytest.mark.parametrize("mode", [1, -1])
def test_oracle_factors_match_brute_force(mode):
    m = 8
    p_min, p_max = (100000, 400000) if mode > 0 else (100000, 104000)
    got = O.sieve_factors(m, p_min, p_max, mode)
    by_p = {}
    for p, b in got:
        by_p.setdefault(int(p), set()).add(int(b))
    primes = [int(p) for p in primes_upto(p_max) if p >= p_min and
              ((mode > 0 and (p - 1) % (3 << m) == 0) or (mode < 0 and p % 10 in (1, 9)))]
    assert primes
    for p in primes:
        assert by_p.get(p, set()) == brute_roots(p, m, mode > 0), p
    assert set(by_p) <= set(primes)


def test_oracle_factors_bmax_and_sizes():
    """Large p: every emitted b is at most bMax and divides; + side counts."""
    m = 12
    got = O.sieve_factors(m, 10 ** 15, 10 ** 15 + 10 ** 10, 1)
    assert len(got) > 0
    for p, b in got[:200]:
        p, b = int(p), int(b)
        assert b <= 5 * 10 ** 9
        x = pow(b, 1 << (m - 1), p)
        assert (x * x - x + 1) % p == 0
