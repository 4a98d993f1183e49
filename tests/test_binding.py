"""The transform binding printed in the integration guide (cyclo's
`transform`, cyclo/src/transform.h:234-655, over include/ctwin.h), compiled
with g++ as a reference-side C++ caller would compile it
(tests/host/cyclo_binding.h), and driven through cyclo::check's PRP loop with
its Gerbicz-Li products (cyclo/src/cyclo.h:252-308, tests/host/cyclo_binding_main.cpp).

- CPU: the binding compiles against the header, links to the library and
  loads (no device call).
- GPU: 64 lanes at n = 2^10 (the reference's VSIZE, at the size of its full-PRP runs):
  RES64 of 2^(N-1) mod N and of its square on sampled lanes against GMP, the
  verdicts, and every lane against the library's own one-call PRP
  (ct_prp_check), which keeps the exponent bits on the device."""
import os
import random
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HOST = os.path.join(HERE, "host")
EXE = os.path.join(HOST, "_build", "cyclo_binding")


def _build():
    r = subprocess.run(["make", "-s", "-C", HOST], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert os.path.exists(EXE)


def test_binding_compiles_links_and_loads():
    _build()
    r = subprocess.run([EXE, "--check"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("ok")


@pytest.mark.gpu
def test_binding_prp_loop_matches_gmp():
    import numpy as np
    ct = pytest.importorskip("ctwin_amd")
    import oracle_lib as O
    if not os.path.exists(EXE):
        _build()
    ln, n = 10, 1 << 10
    rng = random.Random(4242)
    bs = [rng.randrange(1 << 18, 1 << 19) * 2 for _ in range(64)]
    print(f"\n[case] binding PRP loop n=2^10 x64 b[0]={bs[0]}", flush=True)
    r = subprocess.run([EXE, str(ln)] + [str(b) for b in bs], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    rows = [line.split() for line in r.stdout.splitlines()]
    assert len(rows) == 64
    got = {int(l): (int(b), int(p), int(x, 16), int(x2, 16)) for l, b, p, x, x2 in rows}
    for lane in (0, 1, 31, 32, 62, 63):
        b, prm, r64, r64sq = got[lane]
        assert b == bs[lane]
        N = O.modulus(n, b)
        want = O.powm(2, N - 1, N)
        assert r64 == want % (1 << 64), lane
        assert r64sq == (want * want % N) % (1 << 64), lane
        assert prm == int(want == 1)
    t = ct.Transform(ln, 64, ct.KIND_AUTO)
    prm, res64, gl, _, rc = ct.prp_check(t, bs)
    assert rc == 0 and gl.all()
    assert [got[l][2] for l in range(64)] == [int(x) for x in res64]
    assert [got[l][1] for l in range(64)] == [int(bool(p)) for p in prm]
