"""Serving loop at its large classes: n = 2^17 .. 2^20 digits beside
2^12 .. 2^16, 1000 squarings per request, b of 16 .. 22 bits.  Each result is
the same in the mix as alone, and the eight fixture requests
(tests/golden/serve_n2e17_2e20.json, GMP) give GMP's RES64 in the mix; the
same requests run on a standalone autotuned engine give GMP's residue mod
three 31-bit primes as well (every digit)."""
import json
import os
import random

import pytest

pytestmark = pytest.mark.gpu
ct = pytest.importorskip("ctwin_amd")
from ctwin_amd.exponent import top_bits

HERE = os.path.dirname(os.path.abspath(__file__))


def fixture():
    return json.load(open(os.path.join(HERE, "golden", "serve_n2e17_2e20.json")))


def drain(s, want):
    out = {}
    guard = 0
    while len(out) < want:
        s.step()
        for rid, r64, h, lat in s.poll():
            out[rid] = (r64, h, lat)
        guard += 1
        assert guard < 200000
    return out


def requests(FIX):
    reqs = [(r["id"], r["log2_n"], r["b"], r["steps"]) for r in FIX["requests"]]
    rng = random.Random(77)
    for i in range(10):
        ln = rng.randrange(12, 21)
        bits = rng.randrange(16, 23)
        reqs.append((i, ln, rng.randrange(1 << (bits - 1), 1 << bits), 1000))
    rng.shuffle(reqs)
    return reqs


def test_serve_large_classes_mixed_alone_and_gmp():
    FIX = fixture()
    reqs = requests(FIX)
    print(f"\n[case] serve large classes: {[(r[1], r[2]) for r in reqs]}", flush=True)
    s = ct.Server(hbm_budget=24 << 30, slots=2, chunk=125, max_steps=1000)
    mixed = {}
    pending = list(reqs)
    while pending:
        for _ in range(3):
            if pending:
                s.submit(*pending.pop(0))
        s.step()
        for rid, r64, h, lat in s.poll():
            mixed[rid] = (r64, h)
    for rid, v in drain(s, len(reqs) - len(mixed)).items():
        mixed[rid] = v[:2]
    s.close()
    assert len(mixed) == len(reqs)
    for r in FIX["requests"]:
        assert "%016X" % mixed[r["id"]][0] == r["res64"], f"request {r['id']} (n=2^{r['log2_n']}, b={r['b']}) vs GMP"
    print("[serve] fixture requests match GMP in the mix", flush=True)
    alone = ct.Server(hbm_budget=24 << 30, slots=1, chunk=125, max_steps=1000)
    for r in reqs:
        alone.submit(*r)
        (rid, v), = drain(alone, 1).items()
        assert mixed[rid] == v[:2], f"request {r} differs alone vs mixed"
    alone.close()
    print("[serve] every request equal alone and mixed", flush=True)


@pytest.mark.parametrize("idx", range(8))
def test_large_class_engine_residue_mod_q(idx):
    FIX = fixture()
    r = FIX["requests"][idx]
    ln, b, steps = r["log2_n"], r["b"], r["steps"]
    n = 1 << ln
    print(f"\n[case] engine n=2^{ln} b={b} steps={steps}", flush=True)
    w, s = top_bits(b, n, steps + 64)
    i0 = s + steps + 63
    assert i0 == r["i0"]
    t = ct.Transform(ln, 1, ct.KIND_AUTO, csize=0)
    t.set_b([b]); t.set_exponents([w << s]); t.set(1)
    t.run(i0, i0 - steps + 1)
    r64, mods = ct.residue_checks(t.get_int(0), [b], FIX["q"])
    t.close()
    assert "%016X" % r64[0] == r["res64"] and mods[0] == r["mod_q"]


def minus_fixture():
    return json.load(open(os.path.join(HERE, "golden", "serve_minus_n2e17_2e20.json")))


def test_serve_minus_large_classes_mixed_alone_and_gmp():
    """The other twin member at the large classes (transforms of 2^18 .. 2^21
    points): its requests, mixed with N+ requests of the same sizes, give
    GMP's RES64 (tests/golden/serve_minus_n2e17_2e20.json) and the same
    results alone."""
    FIX = minus_fixture()
    reqs = [(r["id"], r["log2_n"], r["b"], r["steps"], -1) for r in FIX["requests"]]
    rng = random.Random(78)
    for i in range(4):
        ln = rng.randrange(17, 21)
        reqs.append((i, ln, rng.randrange(1 << 17, 1 << 20) | 1, 1000, 1))
    rng.shuffle(reqs)
    print(f"\n[case] serve N- large classes: {[(r[1], r[2], r[4]) for r in reqs]}", flush=True)
    s = ct.Server(hbm_budget=24 << 30, slots=2, chunk=125, max_steps=1000)
    for r in reqs:
        s.submit(*r)
    mixed = {rid: v[:2] for rid, v in drain(s, len(reqs)).items()}
    s.close()
    for r in FIX["requests"]:
        assert "%016X" % mixed[r["id"]][0] == r["res64"], f"N- request {r['id']} (n=2^{r['log2_n']}, b={r['b']}) vs GMP"
    alone = ct.Server(hbm_budget=24 << 30, slots=1, chunk=125, max_steps=1000)
    for r in reqs:
        alone.submit(*r)
        (rid, v), = drain(alone, 1).items()
        assert mixed[rid] == v[:2], f"request {r} differs alone vs mixed"
    alone.close()


@pytest.mark.parametrize("idx", range(4))
def test_minus_large_class_engine_residue_mod_q(idx):
    FIX = minus_fixture()
    r = FIX["requests"][idx]
    ln, b, steps = r["log2_n"], r["b"], r["steps"]
    n = 1 << ln
    print(f"\n[case] N- engine n=2^{ln} b={b} steps={steps}", flush=True)
    w, s = top_bits(b, n, steps + 64, sign=-1)
    i0 = s + steps + 63
    assert i0 == r["i0"]
    t = ct.Transform(ln, 1, ct.KIND_AUTO, csize=0, sign=-1)
    t.set_b([b]); t.set_exponents([w << s]); t.set(1)
    t.run(i0, i0 - steps + 1)
    r64, mods = ct.residue_checks(t.get_int(0), [b], FIX["q"])
    t.close()
    assert "%016X" % r64[0] == r["res64"] and mods[0] == r["mod_q"]
