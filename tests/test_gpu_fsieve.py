"""GPU factor emission driven block by block (gfsieve::check,
sieve/GPU/src/ctsieve.h:326-431).

- The block cursor over a p range gives the pairs of the one-call emission
  (whose pairs equal the oracle's, tests/test_gpu_factors.py).
- Completeness at 10^12 .. 10^16 by closed forms, with every root emitted:
  each prime p = 3k 2^m + 1 in the range (primes listed independently by
  sympy) has exactly 2^m roots of N+ = Phi_{3 2^m}(b) mod p; each prime
  p = 10k +- 1 has 0, 2^e or 2^(e+1) roots of N- (e = min(v2(p - 1), m - 1)):
  2^e for each y = (1 +- sqrt 5)/2 with y^((p - 1) / 2^e) = 1.  Every pair is
  also re-checked on the host.
- The file driver writes the reference's fp<ext> / fn<ext> lines and its
  ctx<ext> / ctx_n<ext> checkpoint, and an interrupted run resumed from the
  checkpoint writes the same file as an uninterrupted one."""
import os
from collections import Counter

import numpy as np
import pytest
from sympy import isprime, sqrt_mod

pytestmark = pytest.mark.gpu
ct = pytest.importorskip("ctwin_amd")


@pytest.mark.parametrize("m,p_min,p_max,mode", [
    (12, 10 ** 15, 10 ** 15 + 10 ** 10, 1),
    (12, 10 ** 12, 10 ** 12 + 5 * 10 ** 6, -1),
    (8, 100000, 400000, 1),
])
def test_cursor_equals_one_call(m, p_min, p_max, mode):
    print(f"\n[case] cursor m={m} p=[{p_min}, {p_max}] mode={mode}", flush=True)
    fs = ct.FactorSieve(m, p_min, p_max, mode)
    i0, i1 = fs.blocks()
    parts = [fs.run(i, 1) for i in range(i0, i1)]        # block by block, as the reference reads them
    fs.close()
    got = np.concatenate(parts) if parts else np.zeros((0, 2), np.uint64)
    got = got[np.lexsort((got[:, 1], got[:, 0]))]
    want = ct.sieve_factors(m, p_min, p_max, mode)
    assert np.array_equal(got, want)


def _v2(x):
    v = 0
    while x % 2 == 0:
        x //= 2; v += 1
    return v


@pytest.mark.parametrize("m,p_min,width", [(12, 10 ** 15, 2 * 10 ** 8), (8, 10 ** 16, 10 ** 8), (12, 10 ** 12, 10 ** 8)])
def test_plus_root_counts_closed_form(m, p_min, width):
    p_max = p_min + width
    print(f"\n[case] + roots m={m} p=[{p_min}, {p_max}]", flush=True)
    fs = ct.FactorSieve(m, p_min, p_max, 1)
    fs.set_bmax((1 << 64) - 1)
    i0, i1 = fs.blocks()
    pairs = fs.run(i0, i1 - i0, cap=1 << 24)
    fs.close()
    q = 3 << m
    primes = [p for p in range(p_min + (1 - p_min) % q, p_max + 1, q) if isprime(p)]
    counts = Counter(int(p) for p in pairs[:, 0])
    assert sorted(counts) == primes
    assert set(counts.values()) == {1 << m}
    assert len(set(map(tuple, pairs.tolist()))) == len(pairs)                 # distinct roots
    for p, b in pairs[:: max(1, len(pairs) // 2000)]:
        p, b = int(p), int(b)
        x = pow(b, 1 << (m - 1), p)
        assert (x * x - x + 1) % p == 0
    print(f"[+] {len(primes)} primes x 2^{m} roots", flush=True)


@pytest.mark.parametrize("m,p_min,width", [(12, 10 ** 12, 2 * 10 ** 6), (10, 10 ** 13, 2 * 10 ** 6)])
def test_minus_root_counts_closed_form(m, p_min, width):
    p_max = p_min + width
    print(f"\n[case] - roots m={m} p=[{p_min}, {p_max}]", flush=True)
    fs = ct.FactorSieve(m, p_min, p_max, -1)
    fs.set_bmax((1 << 64) - 1)
    i0, i1 = fs.blocks()
    pairs = fs.run(i0, i1 - i0, cap=1 << 24)
    fs.close()
    counts = Counter(int(p) for p in pairs[:, 0])
    want = {}
    k_lo = p_min // 10
    for k in range(k_lo, p_max // 10 + 2):
        for p in (10 * k - 1, 10 * k + 1):
            if p < p_min or p > p_max or not isprime(p):
                continue
            s5 = sqrt_mod(5, p)
            if s5 is None:
                continue
            e = min(_v2(p - 1), m - 1)
            inv2 = (p + 1) // 2
            c = 0
            for y in ((1 - s5) * inv2 % p, (1 + s5) * inv2 % p):
                if pow(y, (p - 1) >> e, p) == 1:
                    c += 1 << e
            if c:
                want[p] = c
    assert dict(counts) == want
    for p, b in pairs[:: max(1, len(pairs) // 2000)]:
        p, b = int(p), int(b)
        x = pow(b, 1 << (m - 1), p)
        assert (x * x - x - 1) % p == 0
    print(f"[-] {len(want)} primes with roots, {len(pairs)} roots", flush=True)


def test_driver_files_checkpoint_and_resume(tmp_path):
    m, p_min, p_max, mode = 8, 10 ** 15, 10 ** 15 + 5 * 10 ** 9, 1
    ext = f"{m}_{p_min // 10 ** 12}_{p_max // 10 ** 12}.txt"
    print(f"\n[case] driver m={m} p=[{p_min}, {p_max}] ext={ext}", flush=True)
    a = tmp_path / "a"; a.mkdir()
    rc, st = ct.fsieve_check(m, p_min, p_max, mode, str(a))
    assert rc == 0 and not st["resumed"] and st["blocks_done"] == st["blocks_total"] >= 3
    lines = (a / ("fp" + ext)).read_text().splitlines()
    want = ct.sieve_factors(m, p_min, p_max, mode)
    assert lines == [f"{int(p)} {int(b)}" for p, b in want]
    assert (a / ("ctx" + ext)).read_text().split()[0] == str(st["blocks_done"])
    b = tmp_path / "b"; b.mkdir()
    rc, st1 = ct.fsieve_check(m, p_min, p_max, mode, str(b), max_blocks=1)
    assert rc == ct.INTERRUPTED and st1["interrupted"] and st1["blocks_done"] == 1
    assert (b / ("ctx" + ext)).read_text().split()[0] == "1"
    rc, st2 = ct.fsieve_check(m, p_min, p_max, mode, str(b))
    assert rc == 0 and st2["resumed"] and st2["blocks_done"] == st["blocks_done"]
    got = (b / ("fp" + ext)).read_text().splitlines()
    assert sorted(got, key=lambda s: tuple(map(int, s.split()))) == lines
    # the - side names its files fn / ctx_n
    c = tmp_path / "c"; c.mkdir()
    rc, _ = ct.fsieve_check(12, 10 ** 12, 10 ** 12 + 10 ** 6, -1, str(c))
    assert rc == 0 and (c / "fn12_1_1.txt").exists() and (c / "ctx_n12_1_1.txt").exists()
