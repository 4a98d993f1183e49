"""Multipass engine with column pass C fused into the next pass A (one launch
per squaring, the column workgroups exchanging chunk carries through device
memory): the same digits as three launches per squaring, run after run, with
Gerbicz-Li products between runs, for both column tiles, replayed from
captured graphs and launched eagerly, with workgroups named by blockIdx or
ordered by ticket."""
import random

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
ct = pytest.importorskip("ctwin_amd")


def _exponents(bs, n):
    if n <= 1 << 16:
        return [b ** n - b ** (n // 2) for b in bs]
    return [top << s for top, s in (O.top_bits_of_exponent(b, n, 512) for b in bs)]


@pytest.mark.parametrize("ln,lanes,tile,graphs,mode,seq,split", [
    (13, 1, 16, True, 1, False, False), (13, 1, 16, False, 1, True, True), (16, 2, 16, True, 1, False, True),
    (16, 3, 32, False, 1, True, False), (16, 3, 16, True, 2, False, True), (21, 1, 16, True, 1, False, False),
    (21, 1, 32, True, 1, False, False), (21, 1, 32, False, 1, False, True), (21, 1, 32, True, 2, False, True),
    (21, 1, 32, True, 1, True, True), (18, 2, 16, True, 1, 2, True), (21, 1, 32, True, 1, 2, True),
    (21, 1, 32, False, 1, 2, False)])
def test_fused_equals_three_pass(ln, lanes, tile, graphs, mode, seq, split):
    n = 1 << ln
    rng = random.Random(1500 + ln + tile)
    bs = [rng.randrange(1 << 16, 1 << 18) for _ in range(lanes)]
    print(f"\n[case] fused ln={ln} lanes={lanes} tile={tile} graphs={graphs} mode={mode} seq={seq} split={split} bs={bs}", flush=True)
    E = _exponents(bs, n)
    i0 = max(e.bit_length() for e in E) - 1
    engines = []
    for fused in (mode, 0):
        t = ct.Transform(ln, lanes, ct.KIND_PLAIN)
        t.set_column_tile(tile)
        t.set_fused(fused)
        t.set_graphs(graphs)
        t.set_row_form(int(seq) if fused else 0)   # the reference engine keeps one workgroup per row and prime
        t.set_column_split(split and fused)     # and one prime per thread
        t.set_b(bs); t.set_exponents(E); t.set(1); t.copy(1, 0)
        engines.append(t)
    i = i0
    for rep, (steps, gl) in enumerate([(40, 0), (40, 0), (77, 25), (3, 0)]):
        for t in engines:
            t.run(i, i - steps + 1, gl)
        i -= steps
        assert np.array_equal(engines[0].read_state(0), engines[1].read_state(0)), f"run {rep}"
        assert np.array_equal(engines[0].read_state(1), engines[1].read_state(1)), f"run {rep}"
    if ln == 13:       # and the oracle's digits of x after the 160 PRP-loop squarings
        for l, b in enumerate(bs):
            assert np.array_equal(engines[0].read_state(0)[l], O.Engine(n, b, 16).prp_steps(i0 - i, i0))


def test_fused_survives_tile_changes():
    """A fused engine whose column tile changes between runs (the launch
    counters of the other grid shape are cleared) keeps the three-pass digits."""
    ln, lanes = 15, 2
    n = 1 << ln
    rng = random.Random(1600)
    bs = [rng.randrange(1 << 16, 1 << 18) for _ in range(lanes)]
    print(f"\n[case] fused tile changes ln={ln} lanes={lanes} bs={bs}", flush=True)
    E = _exponents(bs, n)
    i = max(e.bit_length() for e in E) - 1
    fused = ct.Transform(ln, lanes, ct.KIND_PLAIN)
    fused.set_fused(True)
    plain = ct.Transform(ln, lanes, ct.KIND_PLAIN)
    for t in (fused, plain):
        t.set_b(bs); t.set_exponents(E); t.set(1)
    for rep, tile in enumerate((16, 32, 16, 32)):
        fused.set_column_tile(tile)
        for t in (fused, plain):
            t.run(i, i - 29)
        i -= 30
        assert np.array_equal(fused.read_state(0), plain.read_state(0)), f"run {rep} tile {tile}"
