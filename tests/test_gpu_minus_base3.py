"""N- = b^n - b^(n/2) - 1 engines and the base-3 option (no engine in the
reference: pinned by Python big integers, as GMP would).  Values are checked,
not digits: the carried digits of these modes have no reference to match."""
import random

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
ct = pytest.importorskip("ctwin_amd")


def modulus(n, b, sign):
    return b ** n - b ** (n // 2) + sign


def value(digits, b):
    return O.digits_value(digits, b)


def rand_state(rng, lanes, n, bs):
    return np.array([[rng.randrange(-b + 1, b) for _ in range(n)] for b in bs], np.int32)


# (log2 n, kind): resident up to a 2^14 transform (n <= 2^13 for N-), multipass from 2^11
MINUS_CASES = [(5, 1), (8, 1), (10, 1), (13, 1), (10, 2), (12, 2), (15, 2)]


@pytest.mark.parametrize("ln,kind", MINUS_CASES)
def test_minus_square_dup_values(ln, kind):
    n = 1 << ln
    rng = random.Random(1100 + ln + 50 * kind)
    lanes = 2 if ln <= 16 else 1
    bs = [rng.randrange(1 << 16, 1 << 20) for _ in range(lanes)]
    print(f"\n[case] N- square_dup ln={ln} kind={kind} bs={bs}", flush=True)
    t = ct.Transform(ln, lanes, kind, sign=-1)
    assert t.kind == kind
    t.set_b(bs)
    x = rand_state(rng, lanes, n, bs)
    t.write_state(0, x)
    vals = [value(x[l], bs[l]) for l in range(lanes)]
    for it in range(3):
        dup = [rng.randrange(2) for _ in range(lanes)]
        t.square_dup(dup)
        for l in range(lanes):
            N = modulus(n, bs[l], -1)
            vals[l] = (vals[l] * vals[l] << dup[l]) % N
    got = t.read_state(0)
    for l in range(lanes):
        N = modulus(n, bs[l], -1)
        assert value(got[l], bs[l]) % N == vals[l]
        assert np.abs(got[l]).max() <= 2 * bs[l]
    canon = t.get_int(0)
    for l in range(lanes):
        assert value(canon[l], bs[l]) == vals[l] and canon[l].max() < bs[l]


@pytest.mark.parametrize("ln,kind", [(6, 1), (12, 1), (11, 2), (15, 2)])
def test_minus_mul_values(ln, kind):
    n = 1 << ln
    rng = random.Random(1200 + ln)
    bs = [rng.randrange(1 << 16, 1 << 20)]
    print(f"\n[case] N- mul ln={ln} kind={kind}", flush=True)
    t = ct.Transform(ln, 1, kind, sign=-1)
    t.set_b(bs)
    x, y = rand_state(rng, 1, n, bs), rand_state(rng, 1, n, bs)
    t.write_state(0, x); t.write_state(1, y)
    t.mul(2, 0, 1)
    N = modulus(n, bs[0], -1)
    assert value(t.read_state(2)[0], bs[0]) % N == value(x[0], bs[0]) * value(y[0], bs[0]) % N


def test_minus_large_tiles_agree():
    """n = 2^18 digits (transform 2^19): the two column tiles give identical
    digits after a run of squarings from a random state (the value check above
    covers the arithmetic; Python's quadratic division makes it slow here)."""
    ln, n = 18, 1 << 18
    rng = random.Random(1400)
    bs = [rng.randrange(1 << 16, 1 << 18)]
    x = rand_state(rng, 1, n, bs)
    print(f"\n[case] N- tiles ln=18 b={bs}", flush=True)
    out = []
    for w in (16, 32):
        t = ct.Transform(ln, 1, ct.KIND_PLAIN, sign=-1)
        t.set_column_tile(w)
        t.set_b(bs)
        t.write_state(0, x)
        for d in (1, 0, 1, 1):
            t.square_dup([d])
        out.append(t.read_state(0))
    assert np.array_equal(out[0], out[1])


def test_minus_resident_and_multipass_agree():
    ln, n = 10, 1 << 10
    rng = random.Random(1300)
    bs = [rng.randrange(1 << 18, 1 << 19) * 2 for _ in range(3)]
    print(f"\n[case] N- resident vs multipass ln=10 bs={bs}", flush=True)
    E = [modulus(n, b, -1) - 1 for b in bs]
    i0 = max(e.bit_length() for e in E) - 1
    out = []
    for kind in (ct.KIND_RESIDENT, ct.KIND_PLAIN):
        t = ct.Transform(ln, 3, kind, sign=-1)
        t.set_b(bs); t.set_exponents(E); t.set(1); t.copy(1, 0)
        t.run(i0, i0 - 299, 40)
        out.append((t.read_state(0), t.read_state(1)))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    for l, b in enumerate(bs):
        N = modulus(n, b, -1)
        assert value(out[0][0][l], b) % N == pow(2, E[l] >> (i0 - 299), N)


@pytest.mark.parametrize("sign,base,kind", [(-1, 2, 1), (-1, 2, 2), (1, 3, 1), (1, 3, 2), (-1, 3, 1)])
def test_prp_check_n2e10_modes(sign, base, kind):
    """Full-PRP shape at n = 1024 (b even in [2^19, 2^20)): base^(N-1) mod N with
    Gerbicz-Li, for N- and for base 3."""
    n, ln = 1024, 10
    rng = random.Random(20261203)
    bs = [rng.randrange(1 << 18, 1 << 19) * 2 for _ in range(4)]
    print(f"\n[case] prp_check n=1024 sign={sign} base={base} kind={kind}", flush=True)
    t = ct.Transform(ln, len(bs), kind, sign=sign)
    t.set_base(base)
    prm, r64, gl, res, rc = ct.prp_check(t, bs, want_residue=True)
    assert rc == 0 and gl.all()
    for l, b in enumerate(bs):
        N = modulus(n, b, sign)
        want = O.powm(base, N - 1, N)
        assert value(res[l], b) == want
        assert int(r64[l]) == want % (1 << 64)
        assert bool(prm[l]) == (want == 1)


def test_known_probable_primes_both_members_and_bases():
    n, ln = 32, 5
    cases = [(-1, 2, [1101, 1154, 1243, 1100]), (1, 3, [1047, 1079, 1111, 1100])]
    for sign, base, bs in cases:
        print(f"\n[case] known prp n=32 sign={sign} base={base} bs={bs}", flush=True)
        t = ct.Transform(ln, len(bs), ct.KIND_RESIDENT, sign=sign)
        t.set_base(base)
        prm, r64, gl, res, rc = ct.prp_check(t, bs)
        assert rc == 0 and gl.all()
        want = [pow(base, modulus(n, b, sign) - 1, modulus(n, b, sign)) == 1 for b in bs]
        assert list(prm) == want and want[:3] == [True, True, True]
