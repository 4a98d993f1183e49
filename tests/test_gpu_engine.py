"""GPU parity of the squaring engine against the oracle (digit-exact state)
and against Python big-integer arithmetic (canonical residues)."""
import random

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

ct = pytest.importorskip("ctwin_amd")


def rand_state(rng, lanes, n, bs):
    return np.array([[rng.randrange(-b + 1, b) for _ in range(n)] for b in bs], np.int32)


def oracle_engines(n, bs):
    return [O.Engine(n, int(b), 16) for b in bs]


def expected_primes(n, bs, np_mode):
    D = max(bs)
    return 3 if np_mode == 3 or 3 * n * D * D >= (ct.moduli()[0] * ct.moduli()[1]) // 2 else 2


@pytest.mark.parametrize("np_mode", [0, 3])
@pytest.mark.parametrize("ln", [5, 6, 7, 8, 9, 10, 11, 12, 13, 14])
def test_square_dup_digit_exact(ln, np_mode):
    n = 1 << ln
    rng = random.Random(ln)
    lanes = 3
    bs = [rng.randrange(1 << 16, 1 << 22) for _ in range(lanes)]
    print(f"\n[case] square_dup ln={ln} np_mode={np_mode} bs={bs}", flush=True)
    t = ct.Transform(ln, lanes, ct.KIND_RESIDENT)
    t.set_b(bs)
    t.set_primes(np_mode)
    assert t.primes == expected_primes(n, bs, np_mode)
    x = rand_state(rng, lanes, n, bs)
    t.write_state(0, x)
    oes = oracle_engines(n, bs)
    ref = x.copy()
    for it in range(3):
        dup = [rng.randrange(2) for _ in range(lanes)]
        t.square_dup(dup)
        for l in range(lanes):
            row = np.ascontiguousarray(ref[l])
            oes[l].square_dup(row, dup[l])
            ref[l] = row
        got = t.read_state(0)
        assert np.array_equal(got, ref), f"step {it}: first mismatch at {np.argwhere(got != ref)[:4]}"


@pytest.mark.parametrize("ln,halves", [(11, False), (12, False), (13, False), (14, False),
                                       (11, True), (12, True), (13, True)])
def test_two_prime_resident_forms_digit_exact(ln, halves):
    """Both two-prime resident forms (primes pass by pass with the skewed
    workgroup-wide exchange, and primes in workgroup halves) against the
    oracle's digits, step by step."""
    n = 1 << ln
    rng = random.Random(900 + ln)
    lanes = 3
    bs = [rng.randrange(1 << 14, 1 << 16) for _ in range(lanes)]
    print(f"\n[case] two-prime resident ln={ln} halves={halves} bs={bs}", flush=True)
    t = ct.Transform(ln, lanes, ct.KIND_RESIDENT)
    t.set_b(bs)
    assert t.primes == 2
    t.set_prime_halves(halves)
    assert t.tuning()[1] == halves
    x = rand_state(rng, lanes, n, bs)
    t.write_state(0, x)
    oes = oracle_engines(n, bs)
    ref = x.copy()
    for it in range(4):
        dup = [rng.randrange(2) for _ in range(lanes)]
        t.square_dup(dup)
        for l in range(lanes):
            row = np.ascontiguousarray(ref[l])
            oes[l].square_dup(row, dup[l])
            ref[l] = row
        got = t.read_state(0)
        assert np.array_equal(got, ref), f"step {it}: first mismatch at {np.argwhere(got != ref)[:4]}"
    t.close()


@pytest.mark.parametrize("ln", [5, 10, 13, 14])
def test_mul_digit_exact(ln):
    n = 1 << ln
    rng = random.Random(100 + ln)
    lanes = 2
    bs = [rng.randrange(33, 1 << 21) for _ in range(lanes)]
    print(f"\n[case] mul ln={ln} bs={bs}", flush=True)
    t = ct.Transform(ln, lanes, ct.KIND_RESIDENT)
    t.set_b(bs)
    x = rand_state(rng, lanes, n, bs)
    y = rand_state(rng, lanes, n, bs)
    t.write_state(0, x); t.write_state(1, y)
    t.mul(2, 0, 1)
    got = t.read_state(2)
    assert np.array_equal(t.read_state(1), y)
    for l in range(lanes):
        e = O.Engine(n, bs[l], 16)
        row = np.ascontiguousarray(x[l]); e.mul(row, np.ascontiguousarray(y[l]))
        assert np.array_equal(got[l], row)


@pytest.mark.parametrize("ln", [6, 10, 12])
def test_run_matches_oracle_prp_loop(ln):
    n = 1 << ln
    rng = random.Random(7 + ln)
    lanes = 2
    bs = [rng.randrange(1 << 17, 1 << 19) * 2 for _ in range(lanes)]
    print(f"\n[case] run ln={ln} bs={bs}", flush=True)
    E = [b ** n - b ** (n // 2) for b in bs]
    i0 = max(e.bit_length() for e in E) - 1
    steps = 64
    t = ct.Transform(ln, lanes, ct.KIND_RESIDENT)
    t.set_b(bs)
    t.set_exponents(E)
    t.set(1)
    t.run(i0, i0 - steps + 1)
    got = t.read_state(0)
    for l in range(lanes):
        want = O.Engine(n, bs[l], 16).prp_steps(steps, i0)
        assert np.array_equal(got[l], want)
        N = O.modulus(n, bs[l])
        assert O.digits_value(got[l], bs[l]) % N == pow(2, E[l] >> (i0 - steps + 1), N)


@pytest.mark.parametrize("kind", [1, 2])
def test_prp_check_n2e10_batch(kind):
    """Full base-2 PRP with Gerbicz-Li at n = 1024, b even in [2^19, 2^20)."""
    n, ln = 1024, 10
    rng = random.Random(20261203)
    bs = [rng.randrange(1 << 18, 1 << 19) * 2 for _ in range(4)]
    print(f"\n[case] prp_check n=1024 kind={kind} bs={bs}", flush=True)
    t = ct.Transform(ln, len(bs), kind)
    prm, r64, gl, res, rc = ct.prp_check(t, bs, want_residue=True)
    assert rc == 0 and gl.all()
    for l, b in enumerate(bs):
        N = O.modulus(n, b)
        want = O.powm(2, N - 1, N)
        assert O.digits_value(res[l], b) == want
        assert int(r64[l]) == want % (1 << 64)
        assert bool(prm[l]) == (want == 1)


def test_prp_check_multipass_n2e13():
    """Full base-2 PRP with Gerbicz-Li at n = 2^13 (~144k-bit N) on the
    multipass engine -- 256-row columns, graphs, with and without the fused
    column pass, both column tiles -- against the resident engine on the same
    candidates: identical canonical residues, RES64 and verdicts.  (GMP's
    2^(N-1) mod N at this size takes minutes per candidate on one core; the
    resident engine is pinned to the oracle digit by digit at this n and to
    GMP on whole PRP tests at n = 2^10.)"""
    n, ln = 1 << 13, 13
    rng = random.Random(1313)
    bs = [rng.randrange(1 << 16, 1 << 17) * 2 for _ in range(2)]
    print(f"\n[case] prp_check multipass vs resident n=2^13 bs={bs}", flush=True)
    ref = ct.Transform(ln, len(bs), ct.KIND_RESIDENT)
    prm0, r640, gl0, res0, rc0 = ct.prp_check(ref, bs, want_residue=True)
    assert rc0 == 0 and gl0.all()
    for tile, fused in ((16, 1), (32, 1), (32, 0)):
        print(f"[case] tile={tile} fused={fused}", flush=True)
        t = ct.Transform(ln, len(bs), ct.KIND_PLAIN)
        t.set_column_tile(tile)
        t.set_fused(fused)
        prm, r64, gl, res, rc = ct.prp_check(t, bs, want_residue=True)
        assert rc == 0 and gl.all()
        assert np.array_equal(res, res0) and np.array_equal(r64, r640) and np.array_equal(prm, prm0)
    for l, b in enumerate(bs):                         # the residue is canonical: below N, digits in [0, b)
        assert int(res0[l].max()) < b
        assert O.digits_value(res0[l], b) < O.modulus(n, b)


def test_prp_check_known_prp():
    from sympy import isprime
    n, ln = 32, 5
    found = []
    b = 1000
    while len(found) < 2:
        if isprime(O.modulus(n, b)):
            found.append(b)
        b += 2
    bs = found + [found[0] + 2]
    print(f"\n[case] known prp n=32 bs={bs}", flush=True)
    t = ct.Transform(ln, len(bs), ct.KIND_RESIDENT)
    prm, r64, gl, res, rc = ct.prp_check(t, bs)
    assert rc == 0 and gl.all()
    assert list(prm) == [True, True, isprime(O.modulus(n, bs[2]))]


# ---------------- plain (multipass) engine ----------------

@pytest.mark.parametrize("np_mode", [0, 3])
@pytest.mark.parametrize("ln", [10, 11, 12, 13, 14, 16, 19])
def test_plain_square_dup_digit_exact(ln, np_mode):
    n = 1 << ln
    rng = random.Random(300 + ln)
    lanes = 2
    bs = [rng.randrange(1 << 16, 1 << 22) for _ in range(lanes)]
    print(f"\n[case] plain square_dup ln={ln} np_mode={np_mode} bs={bs}", flush=True)
    t = ct.Transform(ln, lanes, ct.KIND_PLAIN)
    assert t.kind == ct.KIND_PLAIN
    t.set_b(bs)
    t.set_primes(np_mode)
    assert t.primes == expected_primes(n, bs, np_mode)
    x = rand_state(rng, lanes, n, bs)
    t.write_state(0, x)
    oes = oracle_engines(n, bs)
    ref = x.copy()
    dups = [[rng.randrange(2) for _ in range(lanes)] for _ in range(3)]
    for it in range(3):
        t.square_dup(dups[it])          # carries of the previous step are applied on load
        for l in range(lanes):
            row = np.ascontiguousarray(ref[l]); oes[l].square_dup(row, dups[it][l]); ref[l] = row
    got = t.read_state(0)
    assert np.array_equal(got, ref), f"mismatch at {np.argwhere(got != ref)[:4]}"


@pytest.mark.parametrize("ln", [10, 15, 18])
def test_plain_mul_digit_exact(ln):
    n = 1 << ln
    rng = random.Random(400 + ln)
    bs = [rng.randrange(33, 1 << 20)]
    print(f"\n[case] plain mul ln={ln} bs={bs}", flush=True)
    t = ct.Transform(ln, 1, ct.KIND_PLAIN)
    t.set_b(bs)
    x = rand_state(rng, 1, n, bs); y = rand_state(rng, 1, n, bs)
    t.write_state(0, x); t.write_state(1, y)
    t.mul(0, 0, 1)
    got = t.read_state(0)
    e = O.Engine(n, bs[0], 16)
    row = np.ascontiguousarray(x[0]); e.mul(row, np.ascontiguousarray(y[0]))
    assert np.array_equal(got[0], row)
    assert np.array_equal(t.read_state(1), y)


def test_plain_equals_resident_run():
    ln, n = 14, 1 << 14
    rng = random.Random(5)
    bs = [rng.randrange(1 << 18, 1 << 19) for _ in range(3)]
    print(f"\n[case] plain vs resident ln=14 bs={bs}", flush=True)
    E = [b ** n - b ** (n // 2) for b in bs]
    i0 = max(e.bit_length() for e in E) - 1
    out = []
    for kind in (ct.KIND_RESIDENT, ct.KIND_PLAIN):
        t = ct.Transform(ln, 3, kind)
        t.set_b(bs); t.set_exponents(E); t.set(1); t.copy(1, 0)
        t.run(i0, i0 - 199, 50)
        out.append((t.read_state(0), t.read_state(1)))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_plain_large_first_steps_match_oracle():
    """n = 2^21 (the single-candidate size): first PRP-loop steps digit-exact."""
    ln, n = 21, 1 << 21
    rng = random.Random(21)
    b = 65536 + 2 * rng.randrange(1 << 14)
    print(f"\n[case] plain ln=21 b={b}", flush=True)
    steps = 40          # x reaches full size after ~22 steps
    top, s = O.top_bits_of_exponent(b, n, 64)
    i0 = s + 63
    E_window = top << s          # E's top 64 bits, lower bits zero (never read)
    t = ct.Transform(ln, 1, ct.KIND_PLAIN)
    t.set_b([b]); t.set_exponents([E_window]); t.set(1)
    t.run(i0, i0 - steps + 1)
    got = t.read_state(0)[0]
    want = O.Engine(n, b, 16).prp_steps_window(top, s, i0, steps)
    assert np.array_equal(got, want)
