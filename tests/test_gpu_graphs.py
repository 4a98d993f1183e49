"""Multipass engine: runs launched as replays of captured HIP graphs give
the same digits as the same runs launched kernel by kernel.  Every replay is
checked: each run call below is one replay of one graph, the same graph is
replayed several times, and the host reads the state (a synchronising copy)
between replays."""
import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ct = pytest.importorskip("ctwin_amd")


def _engines(ln, lanes, bs, E):
    out = []
    for graphs in (True, False):
        t = ct.Transform(ln, lanes, ct.KIND_PLAIN)
        t.set_graphs(graphs)
        t.set_b(bs); t.set_exponents(E); t.set(1); t.copy(1, 0)
        out.append(t)
    return out


@pytest.mark.parametrize("ln,lanes", [(12, 3), (16, 2), (21, 1)])
def test_graph_replays_equal_eager(ln, lanes):
    n = 1 << ln
    rng = random.Random(700 + ln)
    bs = [rng.randrange(1 << 16, 1 << 18) for _ in range(lanes)]
    print(f"\n[case] graphs ln={ln} lanes={lanes} bs={bs}", flush=True)
    E = [b ** n - b ** (n // 2) for b in bs] if ln <= 16 else None
    if E is None:
        # the top 256 bits of E for n = 2^21 (lower bits are never read here)
        import oracle_lib as O
        tops = [O.top_bits_of_exponent(b, n, 256) for b in bs]
        E = [top << s for top, s in tops]
    i0 = max(e.bit_length() for e in E) - 1
    g, eager = _engines(ln, lanes, bs, E)
    i = i0
    for rep in range(4):                     # one replay of the 48-squaring graph per call
        for t in (g, eager):
            t.run(i, i - 47)
        i -= 48
        assert np.array_equal(g.read_state(0), eager.read_state(0)), f"replay {rep}"
    for t in (g, eager):                     # runs split at Gerbicz-Li points: graphs of several lengths
        t.run(i, i - 99, 30)
    assert np.array_equal(g.read_state(0), eager.read_state(0))
    assert np.array_equal(g.read_state(1), eager.read_state(1))
    g.close(); eager.close()
