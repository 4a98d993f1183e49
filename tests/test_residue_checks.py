"""ctwin_amd.residue_checks (RES64 and x mod q of canonical digits, vectorised
over lanes) against Python big integers: random digit vectors, odd and even b,
digits at b - 1 (CPU only)."""
import random

import numpy as np

import ctwin_amd as ct

Q = [2147483629, 2147483587, 2147483579]


def test_residue_checks_match_python_ints():
    rng = random.Random(11)
    for n in (1, 2, 7, 64, 1000, 4096):
        bs = [rng.randrange(33, 1 << 31) for _ in range(3)] + [(1 << 18) * 2, (1 << 31) - 1]
        canon = np.array([[rng.randrange(b) for _ in range(n)] for b in bs], np.uint32)
        canon[-1, :] = bs[-1] - 1
        r64, mods = ct.residue_checks(canon, bs, Q)
        for l, b in enumerate(bs):
            x = sum(int(d) * b ** j for j, d in enumerate(canon[l]))
            assert r64[l] == x % (1 << 64)
            assert mods[l] == [x % q for q in Q]
