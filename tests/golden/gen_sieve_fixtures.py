"""Writes the sieve survivor fixtures from the reference CPU sieve.

The reference program is compiled from its own source by oracle/ref/Makefile
into oracle/_ref/ctsieve_ref (this container only; it needs /root/reference).
Each case records the window, the p bounds and the survivors: small cases
keep the full ascending list (delta-encoded), the full window keeps the count,
the SHA-256 of the cand_*.txt text ("b\\n" per survivor) and the first and
last survivors.

With --files it also keeps the reference's own state and candidate files
(Sieve::write(true)) of one tiny window under tests/golden/sieve_files/, the
byte-format fixtures of the file tests.

With --beyond it writes sieve_beyond_2e32.json: windows sieved from explicit
p_min bounds with p across and above 2^32 (the struck b listed in full).

usage: python tests/golden/gen_sieve_fixtures.py [--full] [--files] [--beyond]
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "ctsieve_ref")

SMALL = [
    {"name": "m8_tiny", "m": 8, "b_min": 2, "b_max": 4000, "p_max_pos": 2000000, "p_max_neg": 100000},
    {"name": "m10_small", "m": 10, "b_min": 1000000, "b_max": 1016383, "p_max_pos": 100000000, "p_max_neg": 10000000},
    {"name": "m12_window64k", "m": 12, "b_min": 1 << 20, "b_max": (1 << 20) + (1 << 16) - 1,
     "p_max_pos": 1000000000, "p_max_neg": 100000000},
]
FILES = {"m": 8, "b_min": 2, "b_max": 3001, "p_max_pos": 100000, "p_max_neg": 100000}
FULL = {"name": "m12_window_2e24", "m": 12, "b_min": 1 << 20, "b_max": (1 << 20) + (1 << 24) - 1,
        "p_max_pos": 1 << 32, "p_max_neg": 1 << 32}
# p ranges that cross and lie above 2^32 (explicit p_min, the resume state)
BEYOND = [
    {"name": "m8_across_2e32", "m": 8, "b_min": 1 << 20, "b_max": (1 << 20) + (1 << 16) - 1,
     "p_min_pos": (1 << 32) - (1 << 25), "p_max_pos": (1 << 32) + (1 << 27),
     "p_min_neg": (1 << 32) - (1 << 25), "p_max_neg": (1 << 32) + (1 << 27)},
    {"name": "m12_above_2e32", "m": 12, "b_min": 3000000, "b_max": 3000000 + (1 << 16) - 1,
     "p_min_pos": 1 << 32, "p_max_pos": (1 << 32) + (1 << 30),
     "p_min_neg": 1 << 32, "p_max_neg": (1 << 32) + (1 << 26)},
]


def run(case):
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle", "ref")])
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "s.txt")
        extra = ["-p", str(case["p_min_pos"]), str(case["p_min_neg"])] if "p_min_pos" in case else []
        subprocess.check_call([REF, str(case["m"]), str(case["b_min"]), str(case["b_max"]),
                               str(case["p_max_pos"]), str(case["p_max_neg"]), out] + extra, cwd=d,
                              stdout=subprocess.DEVNULL)
        text = open(out).read()
    surv = [int(x) for x in text.split()]
    return surv, hashlib.sha256(text.encode()).hexdigest()


def main():
    cases = []
    for c in SMALL:
        surv, h = run(c)
        deltas = [surv[0]] + [b - a for a, b in zip(surv, surv[1:])] if surv else []
        cases.append(dict(c, count=len(surv), sha256=h, deltas=deltas))
        print(c["name"], len(surv), flush=True)
    json.dump({"source": "sieve/CPU/ctsieve.cpp check_pos + check_neg from the ini start",
               "cases": cases}, open(os.path.join(HERE, "sieve_small.json"), "w"))
    if "--files" in sys.argv:
        dst = os.path.join(HERE, "sieve_files")
        os.makedirs(dst, exist_ok=True)
        c = FILES
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle", "ref")])
        with tempfile.TemporaryDirectory() as d:
            subprocess.check_call([REF, str(c["m"]), str(c["b_min"]), str(c["b_max"]), str(c["p_max_pos"]),
                                   str(c["p_max_neg"]), os.path.join(d, "s.txt"), d], stdout=subprocess.DEVNULL)
            stem = f"_{c['m']}_{c['b_min']}_{c['b_max']}"
            for name in ("sv" + stem + ".dat", "cand" + stem + ".txt"):
                with open(os.path.join(d, name), "rb") as f, open(os.path.join(dst, name), "wb") as g:
                    g.write(f.read())
        json.dump(dict(FILES, source="sieve/CPU/ctsieve.cpp Sieve::write(true) after check_pos + check_neg"),
                  open(os.path.join(dst, "case.json"), "w"))
        print("files", stem)
    if "--beyond" in sys.argv:
        cases = []
        for c in BEYOND:
            surv, h = run(c)
            sv = set(surv)
            struck = [b for b in range(c["b_min"], c["b_max"] + 1) if b not in sv]
            cases.append(dict(c, count=len(surv), sha256=h, struck=struck))
            print(c["name"], len(surv), "survivors,", len(struck), "struck", flush=True)
        json.dump({"source": "sieve/CPU/ctsieve.cpp check_pos + check_neg from explicit p_min (p across 2^32)",
                   "cases": cases}, open(os.path.join(HERE, "sieve_beyond_2e32.json"), "w"))
    if "--full" in sys.argv:
        surv, h = run(FULL)
        json.dump(dict(FULL, source="sieve/CPU/ctsieve.cpp check_pos + check_neg from the ini start",
                       count=len(surv), sha256=h, head=surv[:32], tail=surv[-32:]),
                  open(os.path.join(HERE, "sieve_m12_window_2e24.json"), "w"))
        print(FULL["name"], len(surv))


if __name__ == "__main__":
    main()
