"""Edge cases of the squaring engine against the oracle (digit-exact):
the largest transform (n = 2^22, three primes), both ends of the b range
(b = 33 and b = 2^31 - 1 in one engine), a lane count that is not a multiple
of 32 (the dup mask spans two words), and digits at the |d| = b - 1 bound."""
import random

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
ct = pytest.importorskip("ctwin_amd")


def _oracle_steps(x, bs, dups):
    """x[lanes][n] after len(dups) squarings, lane l multiplied by 2 when dups[s][l]."""
    n = x.shape[1]
    out = x.copy()
    for l, b in enumerate(bs):
        e = O.Engine(n, int(b), 16)
        row = np.ascontiguousarray(out[l])
        for d in dups:
            e.square_dup(row, int(d[l]))
        out[l] = row
    return out


def test_largest_transform_n2e22():
    ln, n = 22, 1 << 22
    rng = random.Random(2222)
    b = rng.randrange(1 << 20, 1 << 21)                 # 3 n b^2 > p1 p2 / 2: three primes
    print(f"\n[case] n=2^22 b={b}", flush=True)
    x = np.array([[rng.randrange(-b + 1, b) for _ in range(n)]], np.int32)
    dups = [[1], [0]]
    got = []
    for fused in (0, 1):
        t = ct.Transform(ln, 1, ct.KIND_PLAIN)
        t.set_b([b])
        t.set_fused(fused)
        assert t.primes == 3
        t.write_state(0, x)
        t.set_exponents([0b10 << 0])                  # bits i = 1 (set), 0 (clear): the same dups as above
        t.run(1, 0)
        got.append(t.read_state(0))
    assert np.array_equal(got[0], got[1])
    want = _oracle_steps(x, [b], dups)
    assert np.array_equal(got[0], want), f"first mismatch at {np.argwhere(got[0] != want)[:4]}"


@pytest.mark.parametrize("kind,ln", [(1, 8), (1, 14), (2, 12), (2, 15)])
def test_extreme_b(kind, ln):
    n = 1 << ln
    bs = [33, (1 << 31) - 1]
    rng = random.Random(3300 + ln)
    print(f"\n[case] extreme b kind={kind} ln={ln}", flush=True)
    t = ct.Transform(ln, 2, kind)
    t.set_b(bs)
    assert t.primes == 3
    x = np.array([[rng.randrange(-b + 1, b) for _ in range(n)] for b in bs], np.int32)
    t.write_state(0, x)
    dups = [[1, 0], [0, 1], [1, 1]]
    for d in dups:
        t.square_dup(d)
    got = t.read_state(0)
    want = _oracle_steps(x, bs, dups)
    assert np.array_equal(got, want), f"first mismatch at {np.argwhere(got != want)[:4]}"


def test_digits_at_the_bound():
    """Every digit at +-(b - 1), the largest coefficients the carry admits."""
    ln, n = 13, 1 << 13
    bs = [(1 << 31) - 1, 1 << 20]
    print(f"\n[case] bound digits ln={ln}", flush=True)
    x = np.array([[(b - 1) * (1 if (j // 3) % 2 == 0 else -1) for j in range(n)] for b in bs], np.int64).astype(np.int32)
    for kind in (ct.KIND_RESIDENT, ct.KIND_PLAIN):
        t = ct.Transform(ln, 2, kind)
        t.set_b(bs)
        t.write_state(0, x)
        t.square_dup([1, 1])
        got = t.read_state(0)
        want = _oracle_steps(x, bs, [[1, 1]])
        assert np.array_equal(got, want), f"kind {kind}: first mismatch at {np.argwhere(got != want)[:4]}"


@pytest.mark.parametrize("kind,ln", [(1, 10), (2, 13)])
def test_33_lanes_dup_mask(kind, ln):
    n, lanes = 1 << ln, 33
    rng = random.Random(33 + ln)
    bs = [rng.randrange(1 << 16, 1 << 19) for _ in range(lanes)]
    print(f"\n[case] 33 lanes kind={kind} ln={ln}", flush=True)
    t = ct.Transform(ln, lanes, kind)
    t.set_b(bs)
    x = np.array([[rng.randrange(-b + 1, b) for _ in range(n)] for b in bs], np.int32)
    t.write_state(0, x)
    dups = [[rng.randrange(2) for _ in range(lanes)] for _ in range(2)]
    dups[0][32] = 1; dups[1][32] = 0                  # the lane in the second mask word
    for d in dups:
        t.square_dup(d)
    got = t.read_state(0)
    want = _oracle_steps(x, bs, dups)
    assert np.array_equal(got, want), f"first mismatch at {np.argwhere(got != want)[:4]}"
