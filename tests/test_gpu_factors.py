"""GPU factor emission (the sieve/GPU path): the sorted (p, b) set equals the
oracle's restatement of the reference kernels + host check."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
ct = pytest.importorskip("ctwin_amd")

CASES = [
    (8, 100000, 400000, 1),                     # small p: every root below bMax
    (8, 100000, 104000, -1),
    (12, 10 ** 15, 10 ** 15 + 10 ** 10, 1),      # P range of ctsieveg (+)
    (12, 10 ** 12, 10 ** 12 + 5 * 10 ** 6, -1),  # T range of ctsieveg (-)
    (14, 10 ** 16 - 10 ** 11, 10 ** 16 + 10 ** 11, 1),   # bMax switches at 1e16
]


@pytest.mark.parametrize("m,p_min,p_max,mode", CASES)
def test_gpu_factors_match_oracle(m, p_min, p_max, mode):
    print(f"\n[case] factors m={m} p=[{p_min}, {p_max}] mode={mode}", flush=True)
    got = ct.sieve_factors(m, p_min, p_max, mode)
    want = O.sieve_factors(m, p_min, p_max, mode)
    assert len(got) == len(want)
    assert np.array_equal(got, want)
