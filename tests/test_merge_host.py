"""The ctools merge's file handling, on the host (no pair in the window, so
the device is not touched): the reference's skipping rules for candidates
(tools/ctools.cpp:151-175), pairs outside the window (ctools.cpp:196) and a
candidate file larger than the reader's 16 MiB buffer, so that numbers cross
buffer boundaries."""
import os
import random

import pytest

ct = pytest.importorskip("ctwin_amd")


def test_candidates_rules_and_big_file(tmp_path):
    rng = random.Random(5)
    cand = sorted(rng.sample(range(10 ** 11, 10 ** 12), 1_500_000))
    lines = []
    for i, b in enumerate(cand):
        lines.append(b)
        if i % 1000 == 0:
            lines.append(b)                      # repeated: skipped
        if i % 1001 == 0 and i:
            lines.append(cand[i - 1])            # not above the previous: skipped
    cpath, opath = tmp_path / "cand.txt", tmp_path / "out.txt"
    cpath.write_text("".join(f"{b}\n" for b in lines))
    assert os.path.getsize(cpath) > (1 << 24)
    fp = tmp_path / "fp.txt"
    fp.write_text("".join(f"{1000003 + 6 * k} {k + 1}\n" for k in range(100)) + "5 7\n5 7\n")   # all b below the window
    st = ct.merge_factors(12, 10 ** 11, 10 ** 12, str(cpath), [(str(fp), 1, 0, 1)], str(opath))
    assert st == {"candidates": len(cand), "pairs": 101, "in_window": 0, "factors": 0, "bad": 0,
                  "out_of_range": 0, "remaining": len(cand)}
    assert opath.read_text() == "".join(f"{b}\n" for b in cand)


def test_malformed_inputs(tmp_path):
    cpath, opath = tmp_path / "cand.txt", tmp_path / "out.txt"
    cpath.write_text("3\n99999999999999999999999\n")          # above 2^64 - 1
    with pytest.raises(ct.CtError):
        ct.merge_factors(12, 1, 10, str(cpath), [], str(opath))
    cpath.write_text("3\n4\n")
    fp = tmp_path / "fp.txt"
    fp.write_text("1000003 5\n1000033\n")                     # a line without b
    with pytest.raises(ct.CtError):
        ct.merge_factors(12, 100, 200, str(cpath), [(str(fp), 1, None, None)], str(opath))
