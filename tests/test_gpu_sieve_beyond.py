"""Window sieve with primes at and above 2^32.

- Bit-exact against the reference program (sieve/CPU/ctsieve.cpp, its own
  build, from explicit p_min bounds): two windows whose p ranges cross and lie
  above 2^32 (tests/golden/sieve_beyond_2e32.json, the struck b listed).
- The reference's own ini bounds (p+ up to 3 10^8 2^m + 1, p- up to
  10^10 + 1, ctsieve.cpp:797-801) on a small window: a superset of the
  p < 2^32 result, and every b struck only by a prime above 2^32 has a prime
  witness p | N+-(b), p in (2^32, bound], among the pairs of the factor-emission
  cursor over that p range limited to the window's b (checked on the host);
  those pairs' b are exactly the newly struck b."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ct = pytest.importorskip("ctwin_amd")
HERE = os.path.dirname(os.path.abspath(__file__))


def fixture_cases():
    return json.load(open(os.path.join(HERE, "golden", "sieve_beyond_2e32.json")))["cases"]


@pytest.mark.parametrize("idx", [0, 1])
def test_window_across_2e32_matches_reference(idx):
    c = fixture_cases()[idx]
    print(f"\n[case] {c['name']} m={c['m']} b=[{c['b_min']}, {c['b_max']}] p+ [{c['p_min_pos']}, {c['p_max_pos']}] "
          f"p- [{c['p_min_neg']}, {c['p_max_neg']}]", flush=True)
    st = ct.sieve_window(c["m"], c["b_min"], c["b_max"], c["p_max_pos"], c["p_max_neg"],
                         p_min_pos=c["p_min_pos"], p_min_neg=c["p_min_neg"])
    struck = [int(b) for b in np.nonzero(st)[0] + c["b_min"]]
    assert struck == c["struck"], f"{len(struck)} struck vs {len(c['struck'])} by the reference"
    assert len(st) - len(struck) == c["count"]


def _divides(p, b, m, sign):
    x = pow(b, 1 << (m - 1), p)
    return (x * x - x + sign) % p == 0


def test_window_at_reference_ini_bounds():
    m, b_min, b_max = 12, 1 << 20, (1 << 20) + (1 << 14) - 1
    pp, pn = 3 * (10 ** 8 << m) + 1, 10 ** 10 + 1           # ctsieve.cpp:800-801
    print(f"\n[case] ini bounds m={m} b=[{b_min}, {b_max}] p+ <= {pp} p- <= {pn}", flush=True)
    low = ct.sieve_window(m, b_min, b_max, 1 << 32, 1 << 32)
    full = ct.sieve_window(m, b_min, b_max, pp, pn)
    assert np.all(full >= low)
    new = set(int(b) for b in np.nonzero(full & (low == 0))[0] + b_min)
    print(f"[ini] {int(low.sum())} struck below 2^32, {len(new)} more above", flush=True)
    witnesses = {}
    for mode, p_top, sign in ((1, pp, 1), (-1, pn, -1)):
        fs = ct.FactorSieve(m, 1 << 32, p_top, mode)
        fs.set_bmax(b_max)
        i0, i1 = fs.blocks()
        pairs = fs.run(i0, i1 - i0, cap=1 << 16)
        fs.close()
        for p, b in pairs:
            p, b = int(p), int(b)
            if b >= b_min:
                assert _divides(p, b, m, sign), (p, b, mode)
                witnesses.setdefault(b, (p, sign))
    # the window lies below every such p: a b is struck by p exactly when b is a root mod p
    assert set(witnesses) - set(int(b) for b in np.nonzero(low)[0] + b_min) == new
