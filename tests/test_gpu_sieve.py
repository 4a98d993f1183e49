"""GPU window sieve: survivor lists bit-exact against the reference program's
(tests/golden) and the oracle."""
import hashlib
import json
import os
import time

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
ct = pytest.importorskip("ctwin_amd")
HERE = os.path.dirname(os.path.abspath(__file__))


def small_cases():
    return json.load(open(os.path.join(HERE, "golden", "sieve_small.json")))["cases"]


@pytest.mark.parametrize("idx", [0, 1, 2])
def test_gpu_sieve_matches_reference_small(idx):
    c = small_cases()[idx]
    print(f"\n[case] sieve {c['name']}", flush=True)
    st = ct.sieve_window(c["m"], c["b_min"], c["b_max"], c["p_max_pos"], c["p_max_neg"])
    got = ct.survivors(st, c["b_min"])
    want = np.cumsum(np.array(c["deltas"], np.int64)).astype(np.uint64)
    assert len(got) == c["count"]
    assert np.array_equal(got, want)


@pytest.mark.parametrize("mode", [1, -1])
def test_gpu_sieve_sides_match_oracle(mode):
    m, b_min, b_max = 12, (1 << 20) + 12345, (1 << 20) + 12345 + 50000
    print(f"\n[case] sieve sides mode={mode}", flush=True)
    got = ct.sieve_window(m, b_min, b_max, 200000000, 20000000, mode=mode)
    want = O.sieve_window(m, b_min, b_max, 200000000, 20000000, mode=mode)
    assert np.array_equal(got, want)


def test_gpu_sieve_neg_range_off_segment_bounds():
    """An N- p range [12345671, 6e7]: k starts at 1234567, inside a sieve
    segment, and ends inside another."""
    m, b_min, b_max = 12, (1 << 20) + 777, (1 << 20) + 777 + 60000
    print("\n[case] sieve N- offset p range", flush=True)
    got = ct.sieve_window(m, b_min, b_max, 10 ** 6, 60000000, mode=-1, p_min_neg=12345671)
    want = O.sieve_window(m, b_min, b_max, 10 ** 6, 60000000, mode=-1, p_min_neg=12345671)
    assert want.sum() > 0
    assert np.array_equal(got, want)


def test_gpu_sieve_accumulates_into_struck():
    """struck is read and updated, as the reference's _bsieve."""
    m, b_min, b_max = 10, 5000, 9000
    pre = np.zeros(b_max - b_min + 1, np.uint8); pre[::7] = 1
    got = ct.sieve_window(m, b_min, b_max, 10 ** 7, 10 ** 6, struck=pre.copy())
    ref = O.sieve_window(m, b_min, b_max, 10 ** 7, 10 ** 6)
    assert np.array_equal(got, ref | pre)


def test_gpu_sieve_m12_full_window():
    """m = 12, b in [2^20, 2^20 + 2^24), every p < 2^32 on both sides."""
    path = os.path.join(HERE, "golden", "sieve_m12_window_2e24.json")
    if not os.path.exists(path):
        pytest.skip("m = 12 window fixture not generated")
    c = json.load(open(path))
    print(f"\n[case] sieve m = 12 window", flush=True)
    t0 = time.perf_counter()
    st = ct.sieve_window(c["m"], c["b_min"], c["b_max"], c["p_max_pos"], c["p_max_neg"])
    dt = time.perf_counter() - t0
    got = ct.survivors(st, c["b_min"])
    print(f"[sieve] m = 12 window: {len(got)} survivors in {dt:.3f} s", flush=True)
    assert len(got) == c["count"]
    assert [int(x) for x in got[:32]] == c["head"] and [int(x) for x in got[-32:]] == c["tail"]
    text = "".join(f"{int(b)}\n" for b in got)
    assert hashlib.sha256(text.encode()).hexdigest() == c["sha256"]
