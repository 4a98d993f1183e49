"""The GMP fixture helper tests/golden/gen_batched_residues.c (the source of
the batched, large-candidate and serving fixtures) against Python big-integer
arithmetic at small sizes, for both twin members: RES64 and x mod three
primes at several checkpoints of the PRP loop run from a common top bit
(CPU only; skipped where GMP is not installed)."""
import json
import os
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
Q = [2147483629, 2147483587, 2147483579]


@pytest.fixture(scope="module")
def helper():
    if not shutil.which("gcc") or not os.path.exists("/opt/conda/include/gmp.h"):
        pytest.skip("gcc or GMP headers missing")
    out = os.path.join(tempfile.mkdtemp(prefix="ct_gen_"), "gen_batched")
    subprocess.check_call(["gcc", "-O2", "-I/opt/conda/include", os.path.join(HERE, "golden", "gen_batched_residues.c"),
                           "-L/opt/conda/lib", "-lgmp", "-Wl,-rpath,/opt/conda/lib", "-o", out])
    return out


def python_loop(b, n, sign, i0, checkpoints):
    h = b ** (n // 2)
    E = h * h - h - (0 if sign > 0 else 2)
    N = E + 1
    x, rows = 1, []
    for t in range(1, max(checkpoints) + 1):
        i = i0 - (t - 1)
        x = x * x % N
        if i >= 0 and (E >> i) & 1:
            x = 2 * x % N
        if t in checkpoints:
            rows.append(("%016X" % (x % (1 << 64)), [x % q for q in Q]))
    return rows


@pytest.mark.parametrize("sign", [1, -1])
@pytest.mark.parametrize("ln", [5, 6, 7])
def test_gen_batched_matches_python(helper, ln, sign):
    n = 1 << ln
    bs = [33, 40001, 65537, 1048573]
    ts = [1, 7, 40, 120]
    i0 = max((b ** n).bit_length() for b in bs) + 2          # a common top bit above every E: leading zero bits
    cmd = [helper] + (["-m"] if sign < 0 else []) + [str(ln), str(i0), ",".join(map(str, ts))] + [str(b) for b in bs]
    lines = subprocess.check_output(cmd, text=True).strip().splitlines()
    assert len(lines) == len(bs)
    for line, b in zip(lines, bs):
        row = json.loads(line)
        assert row["b"] == b
        want = python_loop(b, n, sign, i0, ts)
        assert row["res64"] == [w[0] for w in want], (b, sign)
        assert row["mod_q"] == [w[1] for w in want], (b, sign)
