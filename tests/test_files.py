"""The reference's file formats on the host (no device): CRC-32, the sieve
state file sv_*.dat and the candidate list cand_*.txt, checked byte for byte
against files written by the reference program itself (tests/golden/sieve_files,
made by gen_sieve_fixtures.py --files)."""
import json
import os

import numpy as np
import pytest

import oracle_lib as O

ct = pytest.importorskip("ctwin_amd")
HERE = os.path.dirname(os.path.abspath(__file__))
FX = os.path.join(HERE, "golden", "sieve_files")


def _lib_built():
    if not os.path.exists(ct.LIB_PATH):
        pytest.skip("libctwin_amd.so not built")


def test_crc32_known_answers():
    _lib_built()
    assert ct.crc32(b"123456789") == 0xCBF43926          # the CRC-32/IEEE check value
    assert ct.crc32(b"") == 0
    assert ct.crc32(b"456789", ct.crc32(b"123")) == ct.crc32(b"123456789")
    import zlib
    data = bytes(range(256)) * 7
    assert ct.crc32(data) == zlib.crc32(data)


def test_sieve_files_match_the_reference_byte_for_byte(tmp_path):
    _lib_built()
    case = json.load(open(os.path.join(FX, "case.json")))
    m, b_min, b_max = case["m"], case["b_min"], case["b_max"]
    struck, p_pos, p_neg = ct.sieve_load(FX, m, b_min, b_max)
    surv = ct.survivors(struck, b_min)
    cand = [int(x) for x in open(os.path.join(FX, f"cand_{m}_{b_min}_{b_max}.txt")).read().split()]
    assert list(surv) == cand
    # the same window sieved by the oracle gives the same survivors
    want = O.sieve_window(m, b_min, b_max, case["p_max_pos"], case["p_max_neg"])
    assert np.array_equal(struck != 0, want != 0)
    ct.sieve_save(str(tmp_path), m, b_min, b_max, p_pos, p_neg, struck, True)
    for name in (f"sv_{m}_{b_min}_{b_max}.dat", f"cand_{m}_{b_min}_{b_max}.txt"):
        assert open(tmp_path / name, "rb").read() == open(os.path.join(FX, name), "rb").read(), name


def test_sieve_file_rotation_and_corruption(tmp_path):
    _lib_built()
    m, b_min, b_max = 9, 100, 1123
    rng = np.random.RandomState(3)
    s1 = rng.randint(0, 2, b_max - b_min + 1).astype(np.uint8)
    s2 = s1.copy(); s2[::7] = 1
    ct.sieve_save(str(tmp_path), m, b_min, b_max, 1000, 2000, s1, False)
    ct.sieve_save(str(tmp_path), m, b_min, b_max, 3000, 4000, s2, False)
    name = tmp_path / f"sv_{m}_{b_min}_{b_max}.dat"
    old = tmp_path / f"sv_{m}_{b_min}_{b_max}.dat.old"
    got, a, b = ct.sieve_load(str(tmp_path), m, b_min, b_max)
    assert np.array_equal(got, s2) and (a, b) == (3000, 4000)
    assert old.exists() and old.read_bytes() != name.read_bytes()
    raw = bytearray(name.read_bytes())
    raw[30] ^= 1
    name.write_bytes(bytes(raw))
    with pytest.raises(ct.CtError):
        ct.sieve_load(str(tmp_path), m, b_min, b_max)
