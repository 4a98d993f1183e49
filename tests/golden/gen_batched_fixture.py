"""Writes tests/golden/batched_n2e14_x256.json: GMP residues of the batched
workload (256 candidates b^(2^14) - b^(2^13) + 1, b uniform in [2^18, 2^19)
from seed 20261203, the same draw as bench.py) after t PRP-loop squarings, at
the checkpoints T below, in two forms:
  "res64" / "mod_q": the lanes in lockstep from the common top bit i0 (the
      largest over the batch), as one launch of the PRP loop runs them; a lane
      whose exponent is shorter sees leading zero bits and stays at x = 1
      for its first i0 - i0(lane) squarings;
  "aligned_res64" / "aligned_mod_q": every lane from its own top bit i0(lane)
      (the engine runs this with the exponent shifted left by i0 - i0(lane)),
      so that every lane carries a non-trivial residue from the start.

Test infrastructure, run in the build container (GMP from /opt/conda):
    python tests/golden/gen_batched_fixture.py
"""
import json
import os
import random
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SEED = 20261203
LN, LANES = 14, 256
T = [20, 50, 100, 200, 500, 1000, 2000, 3000, 4000, 5000]
Q = [2147483629, 2147483587, 2147483579]


def batch_bs():
    rng = random.Random(SEED)
    return [rng.randrange(1 << 18, 1 << 19) for _ in range(LANES)]


def main():
    gen = os.path.join("/tmp", "ct_gen_batched")
    subprocess.check_call(["gcc", "-O2", "-I/opt/conda/include", os.path.join(HERE, "gen_batched_residues.c"),
                           "-L/opt/conda/lib", "-lgmp", "-Wl,-rpath,/opt/conda/lib", "-o", gen])
    bs = batch_bs()
    n = 1 << LN
    i0 = max((b ** n - b ** (n // 2)).bit_length() for b in bs) - 1
    jobs = int(sys.argv[1]) if len(sys.argv) > 1 else (os.cpu_count() or 4)
    groups = [bs[k::jobs] for k in range(jobs)]

    def run(g):
        out = subprocess.check_output([gen, str(LN), str(i0), ",".join(map(str, T))] + [str(b) for b in g], text=True)
        return [json.loads(l) for l in out.splitlines() if l.strip()]

    def run_aligned(g):
        rows = []
        for b in g:
            own = (b ** n - b ** (n // 2)).bit_length() - 1
            out = subprocess.check_output([gen, str(LN), str(own), ",".join(map(str, T)), str(b)], text=True)
            rows += [json.loads(l) for l in out.splitlines() if l.strip()]
        return rows

    with ThreadPoolExecutor(jobs) as ex:
        rows = {r["b"]: r for part in ex.map(run, groups) for r in part}
        arows = {r["b"]: r for part in ex.map(run_aligned, groups) for r in part}
    out = {"log2_n": LN, "seed": SEED, "i0": i0, "q": Q, "t": T,
           "lanes": [{"b": b, "res64": rows[b]["res64"], "mod_q": rows[b]["mod_q"],
                      "aligned_res64": arows[b]["res64"], "aligned_mod_q": arows[b]["mod_q"]} for b in bs]}
    path = os.path.join(HERE, "batched_n2e14_x256.json")
    with open(path, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(f"wrote {path}: {LANES} lanes, i0 = {i0}, checkpoints {T}")


if __name__ == "__main__":
    main()
