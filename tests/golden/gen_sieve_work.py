"""Counts the arithmetic of the 2^24-wide window sieve (m = 12,
b in [2^20, 2^20 + 2^24), p < 2^32 on both sides) as the reference algorithm
does it (sieve/CPU/ctsieve.cpp check_pos / check_neg / check_root, restated in
oracle/ctwin_oracle.c), and writes tests/golden/sieve_work_m12_window_2e24.json:

- the 64-bit Montgomery products of the candidate tests, the non-residue
  searches, the square roots and the root enumeration;
- the products of check_root's strike verification (b^(2^(m-1)) per strike);
- the survivor count and SHA-256, checked against tests/golden/sieve_m12_window_2e24.json
  so the counted run is the one whose output is pinned.

These counts are the algorithmic work of the window: the bench divides them by
the GPU's window time.  One run takes a few minutes on one core."""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402


def main():
    fx = json.load(open(os.path.join(HERE, "sieve_m12_window_2e24.json")))
    m, b_min, b_max = fx["m"], fx["b_min"], fx["b_max"]
    O.sieve_products(reset=True)
    t0 = time.time()
    struck = O.sieve_window(m, b_min, b_max, fx["p_max_pos"], fx["p_max_neg"])
    dt = time.time() - t0
    mont, verify = O.sieve_products(reset=True)
    surv = (np.nonzero(struck == 0)[0] + b_min).astype(np.uint64)
    sha = hashlib.sha256("".join(f"{int(b)}\n" for b in surv).encode()).hexdigest()
    assert len(surv) == fx["count"], (len(surv), fx["count"])
    out = {"name": fx["name"], "m": m, "b_min": b_min, "b_max": b_max,
           "p_max_pos": fx["p_max_pos"], "p_max_neg": fx["p_max_neg"],
           "montgomery_products": mont, "verify_products": verify,
           "survivors": int(len(surv)), "sha256_matches_fixture": sha == fx["sha256"],
           "oracle_seconds_one_core": round(dt, 1)}
    json.dump(out, open(os.path.join(HERE, "sieve_work_m12_window_2e24.json"), "w"), indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
