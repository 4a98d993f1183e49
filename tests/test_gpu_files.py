"""Checkpoint (cyclo.ctx, the reference's layout) and the batch PRP driver
over a candidate file: results in the reference's wording, equal to Python's
pow; an interrupted run resumes from its checkpoint to the same files."""
import os
import random
import struct

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
ct = pytest.importorskip("ctwin_amd")

REF_P = (4293918721, 4253024257, 4231004161)     # cyclo/src/transform.h:26-28


def _rand_state(rng, lanes, n, bs):
    return np.array([[rng.randrange(-b + 1, b) for _ in range(n)] for b in bs], np.int32)


@pytest.mark.parametrize("ln,lanes,kind", [(10, 3, 1), (14, 2, 2)])
def test_context_round_trip_and_layout(tmp_path, ln, lanes, kind):
    n = 1 << ln
    rng = random.Random(1000 + ln)
    bs = [rng.randrange(1 << 16, 1 << 20) for _ in range(lanes)]
    print(f"\n[case] context ln={ln} lanes={lanes} kind={kind}", flush=True)
    t = ct.Transform(ln, lanes, kind)
    t.set_b(bs)
    x0, x1 = _rand_state(rng, lanes, n, bs), _rand_state(rng, lanes, n, bs)
    t.write_state(0, x0); t.write_state(1, x1)
    path = str(tmp_path / "cyclo.ctx")
    t.context_save(path, 3, 12345, 67.5)
    raw = open(path, "rb").read()
    assert len(raw) == 20 + 2 * lanes * n * 12 + 4
    assert struct.unpack("<iiid", raw[:20]) == (1, 3, 12345, 67.5)
    assert struct.unpack("<I", raw[-4:])[0] == (~ct.crc32(raw[:-4]) & 0xFFFFFFFF) ^ 0xA23777AC
    size = 2 * lanes * n
    x12 = np.frombuffer(raw[20:20 + 8 * size], np.uint32).reshape(2, n, lanes, 2)
    x3 = np.frombuffer(raw[20 + 8 * size:-4], np.uint32).reshape(2, n, lanes)
    for r, x in enumerate((x0, x1)):
        d = x.T.astype(np.int64)                          # [j][lane]
        assert np.array_equal(x3[r], np.where(d < 0, d + REF_P[2], d))
        assert np.array_equal(x12[r][..., 0], np.where(d < 0, d + REF_P[0], d))
        assert np.array_equal(x12[r][..., 1], np.where(d < 0, d + REF_P[1], d))
    u = ct.Transform(ln, lanes, kind)
    u.set_b(bs)
    assert u.context_load(path) == (3, 12345, 67.5)
    assert np.array_equal(u.read_state(0), x0) and np.array_equal(u.read_state(1), x1)
    # a damaged file falls back to the rotated previous one
    t.write_state(0, x1)
    t.context_save(path, 4, 99, 1.0)
    bad = bytearray(open(path, "rb").read()); bad[100] ^= 0xFF
    open(path, "wb").write(bytes(bad))
    assert u.context_load(path) == (3, 12345, 67.5)
    assert np.array_equal(u.read_state(0), x0)


def _expected_lines(ln, bs):
    n = 1 << ln
    lines, prps = [], []
    for i, b in enumerate(bs):
        if i > 0 and b == bs[i - 1]:
            continue
        N = O.modulus(n, b)
        r = O.powm(2, N - 1, N)
        head = f"{b}^{n} - {b}^{n // 2} + 1 is "
        if r == 1:
            lines.append(head + "a probable prime."); prps.append(b)
        else:
            lines.append(head + f"composite. RES64: {r % (1 << 64):016X}")
    return lines, prps


def _batches(bs, lanes):
    out = []
    for k in range(0, len(bs), lanes):
        chunk = bs[k:k + lanes]
        out += chunk + [chunk[-1]] * (lanes - len(chunk))
    return out


def test_driver_results_and_resume(tmp_path):
    ln, lanes = 5, 8
    bs = list(range(1100, 1119))          # holds the probable primes 1111 and 1112
    cand = tmp_path / "cand.txt"
    cand.write_text("".join(f"{b}\n" for b in bs))
    print(f"\n[case] driver ln={ln} lanes={lanes} b={bs[0]}..{bs[-1]}", flush=True)
    want, prps = _expected_lines(ln, _batches(bs, lanes))
    assert prps, "the sample holds probable primes"
    a = tmp_path / "a"; a.mkdir()
    rc, st = ct.cyclo_check_file(ln, str(cand), str(a), lanes=lanes)
    assert rc == 0 and st["candidates"] == len(bs) and st["prps"] == len(prps) and not st["resumed"]
    assert (a / "result_c.txt").read_text().splitlines() == want
    assert [int(x) for x in (a / f"prp_{ln}.txt").read_text().split()] == prps
    assert not (a / "cyclo.ctx").exists()
    # interrupted twice (inside the first and the second batch), then resumed
    b = tmp_path / "b"; b.mkdir()
    rc, st = ct.cyclo_check_file(ln, str(cand), str(b), lanes=lanes, max_steps=100)
    assert rc == ct.INTERRUPTED and (b / "cyclo.ctx").exists() and st["squarings"] == 100
    rc, st = ct.cyclo_check_file(ln, str(cand), str(b), lanes=lanes, max_steps=300)
    assert rc == ct.INTERRUPTED and st["resumed"]
    rc, st = ct.cyclo_check_file(ln, str(cand), str(b), lanes=lanes)
    assert rc == 0 and st["resumed"]
    assert (b / "result_c.txt").read_text() == (a / "result_c.txt").read_text()
    assert (b / f"prp_{ln}.txt").read_text() == (a / f"prp_{ln}.txt").read_text()


def test_driver_n2e10_batch(tmp_path):
    """n = 2^10 digits (the plain PRP configuration), one batch: RES64 lines equal pow."""
    ln = 10
    rng = random.Random(20261203)
    bs = [rng.randrange(1 << 18, 1 << 19) * 2 for _ in range(6)]
    cand = tmp_path / "cand.txt"
    cand.write_text("".join(f"{b}\n" for b in bs))
    print(f"\n[case] driver ln=10 bs={bs}", flush=True)
    rc, st = ct.cyclo_check_file(ln, str(cand), str(tmp_path), lanes=8)
    assert rc == 0
    want, _ = _expected_lines(ln, _batches(bs, 8))
    assert (tmp_path / "result_c.txt").read_text().splitlines() == want
