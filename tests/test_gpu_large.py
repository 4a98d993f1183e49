"""One candidate with n = 2^21 digits, b = 65536 + 2u (b = 73764),
10,000 PRP-loop squarings on the plain engine.  Checked against GMP residues
(RES64 and x mod three primes, every 1000 squarings, written by
tests/golden/gen_large_residues.c) and, over all 10,000 squarings, against the
engine's own three-prime path (different arithmetic, same carried digits) and
its fused column pass (pass C and the next pass A in one launch)."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ct = pytest.importorskip("ctwin_amd")
HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "large_n2e21_b73764.jsonl")


def residue_checks(canon, b, qs):
    """RES64 and x mod q of x = sum canon[j] b^j."""
    d = canon.astype(np.uint64)
    n = len(d)
    pw = np.empty(n, np.uint64); pw[0] = 1
    if n > 1:
        pw[1:] = np.uint64(b)
        np.cumprod(pw, out=pw)                       # b^j mod 2^64 (wrapping)
    r64 = int(np.sum(d * pw, dtype=np.uint64))
    mods = []
    for q in qs:
        pq = np.ones(n, np.uint64)
        step, cur = 1, b % q
        while step < n:                               # pq[j] = b^j mod q by doubling
            pq[step:2 * step] = (pq[:min(step, n - step)] * np.uint64(cur)) % np.uint64(q)
            cur = cur * cur % q
            step *= 2
        prod = (d * pq) % np.uint64(q)
        s = 0
        for k in range(0, n, 1 << 15):
            s += int(np.sum(prod[k:k + (1 << 15)], dtype=np.uint64))
        mods.append(s % q)
    return r64, mods


def test_large_candidate_10000_squarings():
    lines = [json.loads(l) for l in open(FIXTURE) if l.strip()]
    head, checks = lines[0], {c["t"]: c for c in lines[1:]}
    ln, b, i0 = head["log2_n"], head["b"], head["i0"]
    n = 1 << ln
    from ctwin_amd.exponent import top_bits
    steps = 10000
    top, s = top_bits(b, n, steps + 64)
    assert s + steps + 63 == i0
    Ew = top << s
    print(f"\n[case] large n=2^{ln} b={b} steps={steps} checkpoints={sorted(checks)}", flush=True)
    engines = []
    for np_mode, fused in ((0, 0), (3, 0), (0, 1)):
        t = ct.Transform(ln, 1, ct.KIND_PLAIN)
        t.set_b([b]); t.set_primes(np_mode); t.set_exponents([Ew]); t.set(1)
        t.set_fused(fused)
        engines.append(t)
    assert engines[0].primes == 2 and engines[1].primes == 3 and engines[2].primes == 2
    done = 0
    for chunk_end in range(1000, steps + 1, 1000):
        states = []
        for t in engines:
            t.run(i0 - done, i0 - chunk_end + 1)
            states.append(t.read_state(0))
        assert np.array_equal(states[0], states[1]), f"2-prime and 3-prime digits differ at t={chunk_end}"
        assert np.array_equal(states[0], states[2]), f"fused column pass digits differ at t={chunk_end}"
        done = chunk_end
        if chunk_end in checks:
            canon = engines[0].get_int(0)[0]
            r64, mods = residue_checks(canon, b, head["q"])
            want = checks[chunk_end]
            assert "%016X" % r64 == want["res64"], f"RES64 at t={chunk_end}"
            assert mods == want["mod_q"], f"x mod q at t={chunk_end}"
            print(f"[large] t={chunk_end} RES64 {want['res64']} matches GMP", flush=True)
