"""Writes tests/golden/prp_n2e10_b2e19.json: GMP Fermat residues of full PRP tests, the
plain PRP of the candidate pair N+- = b^1024 - b^512 +- 1 (n = 2^10 digits),
b an even integer uniform in [2^19, 2^20) drawn from seed 20261203 + 1, for
1024 such b (the first 64 are a batch of the reference's 64 lanes,
cyclo/src/engine.h:19; b[0] is the plain one-candidate case): RES64 of
base^(N-1) mod N for bases 2 and 3 and both members.

Test infrastructure (the oracle's mpz_powm), run in the build container:
    python tests/golden/gen_prp_n2e10_fixture.py
"""
import json
import multiprocessing
import os
import random
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O                                    # noqa: E402

SEED = 20261203
LN, LANES = 10, 1024


def batch_bs():
    rng = random.Random(SEED + 1)
    return [2 * rng.randrange(1 << 18, 1 << 19) for _ in range(LANES)]


def row_for(b):
    n = 1 << LN
    row = {"b": b}
    for sign, name in ((1, "plus"), (-1, "minus")):
        N = b ** n - b ** (n // 2) + sign
        for base in (2, 3):
            r = O.powm(base, N - 1, N)
            row[f"{name}_base{base}"] = {"res64": "%016X" % (r % (1 << 64)), "prp": r == 1}
    return row


def main():
    path = os.path.join(HERE, "prp_n2e10_b2e19.json")
    # rows already in the file for the same b are kept (about 1 s of GMP per
    # residue here); the others are computed on every CPU
    have = {}
    if os.path.exists(path):
        for row in json.load(open(path))["lanes"]:
            have[row["b"]] = row
    bs = batch_bs()
    todo = [b for b in bs if b not in have]
    with multiprocessing.Pool(os.cpu_count()) as pool:
        for row in pool.imap(row_for, todo, chunksize=4):
            have[row["b"]] = row
    rows = [have[b] for b in bs]
    with open(path, "w") as f:
        json.dump({"log2_n": LN, "seed": SEED + 1, "lanes": rows}, f, separators=(",", ":"))
    print(f"wrote {path}: {len(rows)} b")


if __name__ == "__main__":
    main()
