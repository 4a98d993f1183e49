"""The batched workload at its measured shape: 256 candidates b^(2^14) - b^(2^13) + 1,
b uniform in [2^18, 2^19) from the bench's seed, every lane checked against GMP
(tests/golden/batched_n2e14_x256.json, written by gen_batched_fixture.py):
RES64 and x mod three 31-bit primes of the canonical residue at
t = 20 ... 5000 PRP-loop squarings.

- The autotuned engine (the bench's configuration: resident kernels for 256
  lanes) runs the lockstep batch, as the bench times it, in persistent
  launches of 480 to 1000 squarings.
- The resident kernels run every lane from its own top bit (exponents shifted
  to a common top index), so that all 256 lanes hold non-trivial residues.
- The multipass engine with 256 lanes runs the first 200 squarings."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ct = pytest.importorskip("ctwin_amd")
from ctwin_amd.exponent import exponent

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "batched_n2e14_x256.json")))
LN, N = FIX["log2_n"], 1 << FIX["log2_n"]
BS = [L["b"] for L in FIX["lanes"]]


def _check(t, tcount, aligned):
    canon = t.get_int(0)
    r64, mods = ct.residue_checks(canon, BS, FIX["q"])
    k = FIX["t"].index(tcount)
    key_r, key_m = ("aligned_res64", "aligned_mod_q") if aligned else ("res64", "mod_q")
    bad = [l for l, L in enumerate(FIX["lanes"]) if "%016X" % r64[l] != L[key_r][k] or mods[l] != L[key_m][k]]
    assert not bad, f"t={tcount} aligned={aligned}: {len(bad)} lanes differ from GMP, first {bad[:8]}"
    nontrivial = sum(1 for L in FIX["lanes"] if L[key_r][k] != "0000000000000001")
    print(f"[batched] t={tcount} aligned={aligned}: 256 lanes match GMP ({nontrivial} with x != 1)", flush=True)


def test_batched_lockstep_autotuned_matches_gmp():
    print(f"\n[case] batched n=2^{LN} x256 lockstep, autotuned engine", flush=True)
    i0 = FIX["i0"]
    t = ct.Transform(LN, 256, ct.KIND_AUTO, csize=0)
    print(f"[batched] engine kind {t.kind}, tuning {t.tuning()[:3]}", flush=True)
    t.set_b(BS)
    t.set_exponents([exponent(b, N) for b in BS])
    assert t.primes == 2
    t.set(1)
    done = 0
    for tc in (20, 500, 1000, 2000, 3000, 4000, 5000):
        t.run(i0 - done, i0 - tc + 1)              # one persistent launch for the resident kernels
        done = tc
        _check(t, tc, False)
    t.close()


def test_batched_aligned_resident_matches_gmp():
    print(f"\n[case] batched n=2^{LN} x256 aligned, resident kernels", flush=True)
    E = [exponent(b, N) for b in BS]
    top = max(e.bit_length() for e in E)
    i0 = top - 1
    shifted = [e << (top - e.bit_length()) for e in E]
    t = ct.Transform(LN, 256, ct.KIND_RESIDENT)
    t.set_b(BS)
    t.set_exponents(shifted)
    t.set(1)
    done = 0
    for tc in (50, 500, 1000, 2000, 5000):
        t.run(i0 - done, i0 - tc + 1)
        done = tc
        _check(t, tc, True)
    t.close()


def test_batched_aligned_multipass_matches_gmp():
    print(f"\n[case] batched n=2^{LN} x256 aligned, multipass engine", flush=True)
    E = [exponent(b, N) for b in BS]
    top = max(e.bit_length() for e in E)
    i0 = top - 1
    t = ct.Transform(LN, 256, ct.KIND_PLAIN)
    t.set_b(BS)
    t.set_exponents([e << (top - e.bit_length()) for e in E])
    t.set(1)
    done = 0
    for tc in (20, 100, 200):
        t.run(i0 - done, i0 - tc + 1)
        done = tc
        _check(t, tc, True)
    t.close()
