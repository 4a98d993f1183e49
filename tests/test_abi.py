"""The C-ABI library loads and exports every entry point include/ctwin.h
declares (CPU only: no compute calls)."""
import ctypes
import os
import re

import pytest

import ctwin_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "ctwin.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ct_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    assert "ct_engine_create" in names and "ct_sieve_window" in names and "ct_prp_check" in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(ctwin_amd.LIB_PATH):
        pytest.skip("libctwin_amd.so not built (run python -m ctwin_amd.build)")
    L = ctypes.CDLL(ctwin_amd.LIB_PATH)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing
    assert set(ctwin_amd.EXPORTS) <= set(declared())


def test_errors_without_device_are_status_codes():
    if not os.path.exists(ctwin_amd.LIB_PATH):
        pytest.skip("libctwin_amd.so not built")
    L = ctwin_amd.lib()
    h = ctypes.c_void_p()
    # invalid sizes are rejected before any device call
    assert L.ct_engine_create(0, 40, 1, 16, 0, ctypes.byref(h)) == -4
    assert L.ct_engine_create(0, 10, 0, 16, 0, ctypes.byref(h)) == -1
    assert L.ct_engine_create(0, 10, 1, 8, 0, ctypes.byref(h)) == -4
    assert b"log2_n" in L.ct_last_error() or len(L.ct_last_error()) > 0


def test_moduli_are_ntt_primes_below_2_30():
    """The engine's primes: below 2^30 (lazy butterflies need 4p < 2^32),
    descending (the CRT assumes p1 > p2 > p3), 3 * 2^22 | p - 1 (the transform
    of length 2^22 modulo X^n - X^(n/2) + 1 needs a root of order 3n), prime,
    and large enough that three of them hold 3 n b^2 for n = 2^22, b < 2^31."""
    if not os.path.exists(ctwin_amd.LIB_PATH):
        pytest.skip("libctwin_amd.so not built")
    P = ctwin_amd.moduli()
    assert P == sorted(P, reverse=True) and len(set(P)) == 3
    for p in P:
        assert p < 1 << 30 and (p - 1) % (3 << 22) == 0
        assert all(pow(a, p - 1, p) == 1 for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37))
        d, s = p - 1, 0
        while d % 2 == 0:
            d //= 2; s += 1
        for a in (2, 3, 5, 7):
            x = pow(a, d, p)
            assert x in (1, p - 1) or any(pow(x, 2 ** k, p) == p - 1 for k in range(1, s))
    assert 3 * (1 << 22) * (1 << 31) ** 2 < P[0] * P[1] * P[2] // 2


def test_two_prime_mode_keeps_digits_far_inside_2p():
    """Two-prime mode converts a signed digit d to d + 2p in one instruction,
    which needs |d| < 2p and d + 2p < 4p.  The mode is chosen only when
    3 n D^2 (N+; 4 n D^2 for N-) < p1 p2 / 2, D >= every |digit| and >= b;
    a pending carry leaves digits below 2 D.  For every supported n >= 32 the
    largest admissible D keeps 2 D below 2^28, far inside 2 p_min."""
    if not os.path.exists(ctwin_amd.LIB_PATH):
        pytest.skip("libctwin_amd.so not built")
    P = ctwin_amd.moduli()
    h2 = P[0] * P[1] // 2
    for ln in range(5, 23):
        n = 1 << ln
        d = 1
        while 3 * n * (2 * d) ** 2 < h2:      # the largest D with 3 n D^2 < p1 p2 / 2: doubling, then bisection
            d *= 2
        lo, hi = d, 2 * d                      # 3 n lo^2 < p1 p2 / 2 <= 3 n hi^2
        while hi - lo > 1:
            mid = (lo + hi) // 2
            lo, hi = (mid, hi) if 3 * n * mid * mid < h2 else (lo, mid)
        assert 2 * lo < (1 << 28) < 2 * min(P)


def test_server_submit_checks_arguments_before_any_device_call():
    if not os.path.exists(ctwin_amd.LIB_PATH):
        pytest.skip("libctwin_amd.so not built")
    s = ctwin_amd.Server(hbm_budget=1 << 30, slots=1, chunk=64, max_steps=64)
    with pytest.raises(ctwin_amd.CtError):
        s.submit(0, 12, 40000, 10, 0)          # sign must be +1 or -1
    with pytest.raises(ctwin_amd.CtError):
        s.submit(0, 22, 40000, 10, -1)         # N- up to 2^21 digits
    with pytest.raises(ctwin_amd.CtError):
        s.submit(0, 12, 32, 10, -1)            # b > 32
    with pytest.raises(ctwin_amd.CtError):
        s.submit(0, 12, 40000, 65, 1)          # steps <= max_steps
    s.submit(0, 22, 40000, 10, 1)
    s.submit(1, 21, 40000, 10, -1)
    assert s.stats()["queued"] == 2
    s.close()
