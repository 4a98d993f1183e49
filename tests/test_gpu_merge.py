"""The ctools merge of factor files into a candidate list
(tools/ctools.cpp:151-248), with the pair check on the device.

- Pairs emitted by the GPU factor sieve, written as fp/fn files in the
  reference's "p b" line format (with repeated lines), merged into a
  candidate list that holds their b values and random others: the remaining
  list is the candidates minus every b in the window with a pair for which p
  divides N+- = b^(2^m) - b^(2^(m-1)) +- 1, computed here with Python's pow.
- The reference's rules on hand-made files: candidates that do not increase
  are skipped, a repeated line counts once, a pair that does not divide is
  counted as bad and removes nothing, p outside the file's range is counted
  and still used, b outside the window is ignored, and p above 2^63 (m = 1:
  b a root of b^2 - b +- 1 mod p, from a square root of -3 or 5) verifies."""
import random

import pytest
from sympy import isprime, sqrt_mod

pytestmark = pytest.mark.gpu
ct = pytest.importorskip("ctwin_amd")


def _divides(p, b, m, mode):
    # p | b^(2^m) - b^(2^(m-1)) + mode
    return (pow(b, 1 << m, p) - pow(b, 1 << (m - 1), p) + mode) % p == 0


def _write_pairs(path, pairs, repeat_every=7):
    with open(path, "w") as f:
        for i, (p, b) in enumerate(pairs):
            f.write(f"{p} {b}\n")
            if i % repeat_every == 0:
                f.write(f"{p} {b}\n")          # a repeated line counts once


@pytest.mark.parametrize("m,ranges", [
    (12, [(10 ** 15, 10 ** 15 + 10 ** 10, 1), (10 ** 12, 10 ** 12 + 5 * 10 ** 6, -1)]),
    (8, [(100000, 400000, 1), (100000, 400000, -1)]),
])
def test_merge_emitted_factors(tmp_path, m, ranges):
    print(f"\n[case] merge m={m} ranges={ranges}", flush=True)
    rng = random.Random(31 + m)
    files, allpairs = [], []
    for k, (p_min, p_max, mode) in enumerate(ranges):
        pairs = [(int(p), int(b)) for p, b in ct.sieve_factors(m, p_min, p_max, mode)]
        assert pairs, "the range holds factors"
        rng.shuffle(pairs)
        path = tmp_path / f"{'fp' if mode > 0 else 'fn'}{m}_{k}.txt"
        _write_pairs(path, pairs)
        files.append((str(path), mode, p_min // 10 ** 12, -(-p_max // 10 ** 12)))
        allpairs += [(p, b, mode) for p, b in pairs]
    bs = sorted({b for _, b, _ in allpairs})
    b_min, b_max = bs[len(bs) // 4], bs[-1]                  # some pairs fall below the window
    others = {rng.randrange(b_min, b_max + 1) for _ in range(2000)}
    cand = sorted(set(b for b in bs if b >= b_min) | others)
    cpath, opath = tmp_path / "cand.txt", tmp_path / "out.txt"
    cpath.write_text("".join(f"{b}\n" for b in cand))
    st = ct.merge_factors(m, b_min, b_max, str(cpath), files, str(opath))
    killed = {b for p, b, mode in allpairs if b_min <= b <= b_max and _divides(p, b, m, mode)}
    want = [b for b in cand if b not in killed]
    got = [int(x) for x in opath.read_text().split()]
    assert got == want
    in_window = sum(1 for _, b, _ in allpairs if b_min <= b <= b_max)
    assert st["candidates"] == len(cand) and st["pairs"] == len(allpairs)
    assert st["in_window"] == in_window and st["factors"] == in_window and st["bad"] == 0
    assert st["out_of_range"] == 0 and st["remaining"] == len(want)


def _prime_3mod4(lo, rng, cond):
    while True:
        p = rng.randrange(lo, 1 << 64) | 3
        if cond(p) and isprime(p):
            return p


def test_merge_rules_and_large_p(tmp_path):
    rng = random.Random(77)
    m = 1
    # p above 2^63, p = 3 mod 4: b = (1 +- sqrt(-3)) / 2 is a root of b^2 - b + 1 (p = 1 mod 3),
    # b = (1 +- sqrt 5) / 2 a root of b^2 - b - 1 (p = +-1 mod 5)
    plus, minus = [], []
    for _ in range(3):
        p = _prime_3mod4(1 << 63, rng, lambda q: q % 3 == 1)
        s = sqrt_mod(p - 3, p)
        plus.append((p, (1 + s) * pow(2, -1, p) % p))
        p = _prime_3mod4(1 << 63, rng, lambda q: q % 5 in (1, 4))
        s = sqrt_mod(5, p)
        minus.append((p, (1 + s) * pow(2, -1, p) % p))
    for p, b in plus:
        assert _divides(p, b, m, 1)
    for p, b in minus:
        assert _divides(p, b, m, -1)
    wanted_b = sorted(b for _, b in plus + minus)
    b_min, b_max = wanted_b[0] - 10, wanted_b[-1] + 10
    spare = [b_min + 1, b_min + 2, b_max - 1]
    cand = sorted(set(wanted_b) | set(spare))
    # the candidate file: a repeated entry and a smaller one after a larger are skipped
    lines = [cand[0], cand[1], cand[1], cand[0]] + cand[2:]
    cpath, opath = tmp_path / "cand.txt", tmp_path / "out.txt"
    cpath.write_text("".join(f"{b}\n" for b in lines))
    bad = (plus[0][0], plus[0][1] + 1)                  # does not divide
    fp_lines = [plus[0], plus[0], plus[1], bad, (plus[2][0], 5)] + [plus[2]]    # (p, 5): b outside the window
    fn_lines = minus
    fp, fn = tmp_path / "fp1_x.txt", tmp_path / "fn1_x.txt"
    fp.write_text("".join(f"{p} {b}\n" for p, b in fp_lines))
    fn.write_text("".join(f"{p} {b}\n" for p, b in fn_lines))
    # p range of the fn file in units of 10^12 set below its primes: all three out of range
    files = [(str(fp), 1, 9223372, 18446745), (str(fn), -1, 1, 2)]
    st = ct.merge_factors(m, b_min, b_max, str(cpath), files, str(opath))
    got = [int(x) for x in opath.read_text().split()]
    killed = {b for _, b in plus + minus}
    assert got == [b for b in cand if b not in killed]
    assert st["candidates"] == len(cand)
    assert st["pairs"] == 5 + 3                         # the repeated line of the fp file counts once
    assert st["in_window"] == 4 + 3 and st["factors"] == 3 + 3 and st["bad"] == 1
    assert st["out_of_range"] == 3 and st["remaining"] == len(spare)


def test_merge_rejects_bad_arguments(tmp_path):
    cpath = tmp_path / "cand.txt"
    cpath.write_text("5\n")
    with pytest.raises(ct.CtError):
        ct.merge_factors(8, 1, 10, str(cpath), [(str(tmp_path / "missing.txt"), 1, None, None)], str(tmp_path / "o.txt"))
    with pytest.raises(ct.CtError):
        ct.merge_factors(8, 1, 10, str(tmp_path / "missing.txt"), [], str(tmp_path / "o.txt"))
    with pytest.raises(ct.CtError):
        ct.merge_factors(8, 1, 10, str(cpath), [(str(cpath), 2, None, None)], str(tmp_path / "o.txt"))
