"""ct_sieve_factors_noalloc (no device allocation; scratch in a fixed device
array) gives the pairs of ct_sieve_factors, through the C ABI: two calls in a
row on the same device, the second over a smaller p range, both sides; and a
too-small output buffer reports CT_ERR_NOMEM with the count needed.  The
output buffers are device memory from torch (plumbing only)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ct = pytest.importorskip("ctwin_amd")
torch = pytest.importorskip("torch")


def _noalloc(m, p_min, p_max, mode, cap):
    L = ct.lib()
    f = L.ct_sieve_factors_noalloc
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p,
                  ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64), ctypes.c_void_p]
    f.restype = ctypes.c_int
    buf = torch.zeros(2 * max(cap, 1), dtype=torch.int64, device="cuda:0")
    torch.cuda.synchronize()
    cnt = ctypes.c_uint64()
    rc = f(0, m, p_min, p_max, mode, ctypes.c_void_p(buf.data_ptr()), cap, ctypes.byref(cnt), None)
    out = buf.cpu().numpy().view(np.uint64)
    k = min(cnt.value, cap)
    pairs = out[:2 * k].reshape(-1, 2)
    return rc, cnt.value, pairs[np.lexsort((pairs[:, 1], pairs[:, 0]))]


@pytest.mark.parametrize("m,ranges,mode", [
    (12, [(10 ** 15, 10 ** 15 + 10 ** 10), (10 ** 15 + 10 ** 9, 10 ** 15 + 3 * 10 ** 9)], 1),
    (12, [(10 ** 12, 10 ** 12 + 5 * 10 ** 6), (10 ** 12 + 10 ** 6, 10 ** 12 + 2 * 10 ** 6)], -1),
    (8, [(100000, 400000), (100000, 150000)], 1),
])
def test_noalloc_equals_allocating_entry_point(m, ranges, mode):
    for p_min, p_max in ranges:                       # the second call covers a smaller range
        print(f"\n[case] noalloc factors m={m} p=[{p_min}, {p_max}] mode={mode}", flush=True)
        want = ct.sieve_factors(m, p_min, p_max, mode)
        rc, cnt, got = _noalloc(m, p_min, p_max, mode, len(want) + 16)
        assert rc == 0 and cnt == len(want)
        assert np.array_equal(got, want)


def test_noalloc_small_buffer_reports_size():
    m, p_min, p_max = 8, 100000, 400000
    want = ct.sieve_factors(m, p_min, p_max, 1)
    print(f"\n[case] noalloc small buffer: {len(want)} pairs, cap 10", flush=True)
    rc, cnt, got = _noalloc(m, p_min, p_max, 1, 10)
    assert rc == -3 and cnt == len(want) and len(got) == 10
    assert all(any((row == w).all() for w in want) for row in got)
