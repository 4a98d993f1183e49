"""Autotuner (csize = 0 at engine creation) and the multipass column tile:
every configuration gives the same carried digits."""
import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ct = pytest.importorskip("ctwin_amd")


def _run(t, bs, n, steps, seed):
    rng = random.Random(seed)
    E = [b ** n - b ** (n // 2) for b in bs] if n <= 1 << 16 else None
    if E is None:
        import oracle_lib as O
        E = [top << s for top, s in (O.top_bits_of_exponent(b, n, 128) for b in bs)]
    i0 = max(e.bit_length() for e in E) - 1
    t.set_b(bs); t.set_exponents(E); t.set(1)
    t.run(i0, i0 - steps + 1)
    return t.read_state(0)


@pytest.mark.parametrize("ln,lanes", [(12, 1), (14, 1), (14, 32), (17, 2)])
def test_tuned_engine_matches_untuned(ln, lanes):
    n = 1 << ln
    rng = random.Random(800 + ln + lanes)
    bs = [rng.randrange(1 << 17, 1 << 20) for _ in range(lanes)]
    print(f"\n[case] tuned ln={ln} lanes={lanes}", flush=True)
    tuned = ct.Transform(ln, lanes, ct.KIND_AUTO, csize=0)
    w, split, fused, ms, rs = tuned.tuning()
    assert len(ms) == ct.TUNE_TIMINGS
    timed = [m for m in ms[:11] if m > 0]       # the kind and multipass choices (three primes)
    assert timed, ms
    # the fused column pass is timed exactly for N+ with 256-row columns
    assert (ms[5] > 0) == (ln >= 13 and ms[1] > 0)
    # the chosen configuration is the fastest one timed
    best = min(timed)
    if tuned.kind == ct.KIND_RESIDENT:
        assert ms[0] == best and w == 0 and not fused and not rs
        # two primes in workgroup halves where timed faster (2^10 .. 2^13)
        assert split == (ms[12] > 0 and ms[12] < ms[11])
    else:
        assert w in (16, 32, 64)
        if fused:
            idx = (5 if w == 16 else 6) + (2 if split else 0)
        else:
            idx = (1 if w in (16, 64) else 2) + (2 if split else 0)
        assert ms[8 + rs if rs else idx] == best
    ref = ct.Transform(ln, lanes, ct.KIND_AUTO, csize=16)
    a = _run(tuned, bs, n, 40, 1)
    b = _run(ref, bs, n, 40, 1)
    assert np.array_equal(a, b)
    # set_b moved the engine to tuned.primes primes; a multipass engine is
    # tuned again for that count and runs the fastest configuration timed
    w2, split2, fused2, ms2, rs2 = tuned.tuning()
    if tuned.kind == ct.KIND_PLAIN:
        timed2 = [m for m in ms2[1:11] if m > 0]
        idx2 = ((5 if w2 == 16 else 6) if fused2 else (1 if w2 in (16, 64) else 2)) + (2 if split2 else 0)
        assert ms2[8 + rs2 if rs2 else idx2] == min(timed2)
    # second creation hits the cache: same choice, same timings
    again = ct.Transform(ln, lanes, ct.KIND_AUTO, csize=0)
    assert again.kind == tuned.kind and again.tuning() == (w, split, fused, ms, rs)


@pytest.mark.parametrize("ln", [13, 16, 20])
def test_column_tiles_give_identical_digits(ln):
    n = 1 << ln
    rng = random.Random(900 + ln)
    bs = [rng.randrange(1 << 16, 1 << 18)]
    print(f"\n[case] tiles ln={ln}", flush=True)
    out = []
    for w in (16, 32):
        for split in (False, True):
            for form in (0, 1, 2):
                t = ct.Transform(ln, 1, ct.KIND_PLAIN)
                t.set_column_tile(w)
                t.set_column_split(split)
                t.set_row_form(form)
                assert t.tuning()[:3] == (w, split, False) and t.tuning()[4] == form
                out.append(_run(t, bs, n, 30, 2))
    for o in out[1:]:
        assert np.array_equal(out[0], o)
