"""Writes tests/golden/serve_n2e17_2e20.json: GMP residues of served requests
of the large size classes of the served stream (n = 2^17 .. 2^20 digits, b of
16 .. 22 bits, 1000 PRP-loop squarings each from the request's own top bit,
the result a request returns).  b is odd so that RES64 depends on every digit;
x mod three 31-bit primes is recorded as well.  With --minus it writes
tests/golden/serve_minus_n2e17_2e20.json: the same for requests of the other
twin member, N = b^n - b^(n/2) - 1.

Test infrastructure, run in the build container (GMP from /opt/conda; the
n = 2^20, 22-bit requests take several minutes each):
    python tests/golden/gen_serve_fixture.py [--minus]
"""
import json
import os
import random
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
STEPS = 1000
Q = [2147483629, 2147483587, 2147483579]
# (log2 n, bits of b)
CASES = [(17, 16), (17, 22), (18, 19), (18, 21), (19, 17), (19, 22), (20, 18), (20, 22)]
CASES_MINUS = [(17, 20), (18, 17), (19, 19), (20, 18)]       # transforms of 2^18 .. 2^21 points


def requests(minus=False):
    rng = random.Random(5006 if minus else 5005)
    out = []
    for i, (ln, bits) in enumerate(CASES_MINUS if minus else CASES):
        b = rng.randrange(1 << (bits - 1), 1 << bits) | 1
        row = {"id": (950 if minus else 900) + i, "log2_n": ln, "b": b, "steps": STEPS}
        if minus:
            row["sign"] = -1
        out.append(row)
    return out


def main():
    gen = os.path.join("/tmp", "ct_gen_batched")
    subprocess.check_call(["gcc", "-O2", "-I/opt/conda/include", os.path.join(HERE, "gen_batched_residues.c"),
                           "-L/opt/conda/lib", "-lgmp", "-Wl,-rpath,/opt/conda/lib", "-o", gen])
    minus = "--minus" in sys.argv[1:]
    reqs = requests(minus)

    def run(r):
        n = 1 << r["log2_n"]
        b = r["b"]
        h = b ** (n // 2)
        i0 = (h * h - h - (2 if minus else 0)).bit_length() - 1     # the request's own top bit of E = N - 1
        cmd = [gen] + (["-m"] if minus else []) + [str(r["log2_n"]), str(i0), str(STEPS), str(b)]
        row = json.loads(subprocess.check_output(cmd, text=True))
        return dict(r, i0=i0, res64=row["res64"][0], mod_q=row["mod_q"][0])

    with ThreadPoolExecutor(len(reqs)) as ex:
        rows = list(ex.map(run, reqs))
    path = os.path.join(HERE, "serve_minus_n2e17_2e20.json" if minus else "serve_n2e17_2e20.json")
    with open(path, "w") as f:
        json.dump({"q": Q, "requests": rows}, f, indent=1)
    print(f"wrote {path}: {len(rows)} requests")


if __name__ == "__main__":
    main()
