"""Benchmark of the Fermat squaring engine on one MI355X.

Workload (the batched configuration): 256 candidates N = b^n - b^(n/2) + 1,
n = 2^14 digits, b uniform in [2^18, 2^19) from a fixed seed, run through the
base-2 PRP loop (x <- x^2 * 2^bit, each lane from its own top exponent bit) in
one persistent launch of the autotuned engine; a step is a chunk of 250
squarings of all 256 candidates (the default 20 steps are the configuration's
5000 squarings per candidate, timed as one launch), and every lane is checked
against GMP after the timed run.  Beside it: the plain one-candidate engine
at n = 2^14 (batched / plain ratio) and the single large candidate at n = 2^21
(b = 65536 + 2u), with its HBM roofline, and the oracle timed on the host CPU.

python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import math
import os
import random
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
VALU_PEAK_T = 39.3          # 256 CU x 64 lanes x 2.4 GHz, 32-bit ops (mul_hi/mul_lo full rate on gfx950)
SEED = 20261203
SQ_PER_STEP = 250           # squarings of every batched candidate per bench step
TUNE_NAMES = ("resident", "multipass_w16", "multipass_w32", "multipass_w16_split", "multipass_w32_split",
              "multipass_w16_fused", "multipass_w32_fused", "multipass_w16_fused_split", "multipass_w32_fused_split",
              "multipass_best_row_primes_in_sequence", "multipass_best_row_primes_in_slices",
              "resident_two_primes", "resident_two_primes_in_halves")


def mul32_alg(n, primes=3):
    """Algorithmic 32-bit multiplies per squaring: per prime n (log2 n + 2.5)
    modular products for the transforms, plus the CRT (three primes: 3 n
    products, two primes: n), 3 mul32 per product -- 9 n (log2 n + 4) with
    three primes, 6 n (log2 n + 3) with two."""
    if primes == 3:
        return 9 * n * (math.log2(n) + 4)
    return 6 * n * (math.log2(n) + 3)


# Measured HBM traffic per launch (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE
# on `bench.py --profile`, raw KiB scaled by what the same counters read for a
# copy kernel of known size: FETCH x 1.994, WRITE x 1.000).  The resident
# kernel reads its state and tables once per launch whatever the number of
# squarings; the n = 2^21 figures are per squaring (row pass + fused column pass).
TRAFFIC_RESIDENT_LAUNCH = (24441 * 1.994 + 45056) * 1024     # k_res_steps<14, 2>, 256 lanes
TRAFFIC_N2E21_SQUARING = (33174 * 1.994 + 16384 + 8621 * 1.994 + 16904) * 1024   # k_rowB + k_colCA2
# VALU issue slots (rocprofv3 --pmc SQ_INSTS_VALU SQ_ACTIVE_INST_VALU2 on
# `bench.py --profile` and `--sieve-only`): instructions minus the quad-cycles
# in which gfx950 issues two VALU instructions at once (mostly v_mov_b32 beside
# an arithmetic instruction).  One slot takes a SIMD 4 clocks, so the issue
# ceiling is 1024 SIMDs x 2.4 GHz / 4 = 614.4 G slots/s.
VALU_SLOT_PEAK_G = 1024 * 2.4 / 4
VALU_SLOTS_PER_SQUARING = (901176042 - 18359072) / 49                # k_res_steps<14, 2>, 256 lanes: 50- minus 1-squaring launch
VALU_SLOTS_N2E21_SQUARING = 6942052 + 6458985                   # k_rowB<21, 8> + k_colCA2<21, 8, 32>
SIEVE_ROOTS_NEG_SLOTS = 7505592064 - 1826200320                 # k_roots_neg, m = 12 window of 2^24 b, its 26 launches
# k_roots_neg per launch on that window: the prime list it reads (FETCH x 1.994)
# and the write-back of the struck byte map (WRITE), both in HBM bytes
TRAFFIC_SIEVE_ROOTS_NEG_LAUNCH = (401214.9 / 52 * 1.994 + 750543.7 / 52) * 1024


def run_prp_n2e10(ct):
    """Full Fermat PRP tests with Gerbicz-Li at n = 2^10 of the candidate
    pair N+- = b^1024 - b^512 +- 1, b even, uniform in [2^19, 2^20) (seed +
    1): base 3 (the configuration's) for both members and base 2 (the
    reference's) for N+, as a batch of 64 (the reference's lane count), as
    a batch of 1024 (four workgroups per CU) and as the plain one-candidate
    engine; RES64 of every test against GMP (tests/golden/prp_n2e10_b2e19.json)."""
    ln, wide, lanes = 10, 1024, 64
    rng = random.Random(SEED + 1)
    bs = [2 * rng.randrange(1 << 18, 1 << 19) for _ in range(wide)]
    fx_path = os.path.join(ROOT, "tests", "golden", "prp_n2e10_b2e19.json")
    fx = json.load(open(fx_path)) if os.path.exists(fx_path) else None
    if fx and [L["b"] for L in fx["lanes"]] != bs:
        fx = None
    out = {"log2_n": ln, "b_range": [1 << 19, 1 << 20], "seed": SEED + 1}
    bad = 0
    checked = 0
    for name, sign, base in (("plus_base3", 1, 3), ("minus_base3", -1, 3), ("plus_base2", 1, 2)):
        for count in (wide, lanes, 1):
            t = ct.Transform(ln, count, ct.KIND_AUTO, csize=0, sign=sign)
            t.set_base(base)
            ct.prp_check(t, bs[:count])                         # warm-up (tuned engine, first launches)
            t0 = time.perf_counter()
            prm, r64, gl, _, rc = ct.prp_check(t, bs[:count])
            dt = time.perf_counter() - t0
            t.close()
            key = f"{name}_{'plain1' if count == 1 else f'batched{count}'}"
            out[key] = {"seconds_per_batch": dt, "candidates_per_hour": 3600.0 * count / dt,
                        "gerbicz_ok": bool(gl.all()), "status": int(rc)}
            if fx:
                for l in range(count):
                    checked += 1
                    bad += ("%016X" % int(r64[l])) != fx["lanes"][l][name]["res64"]
    out["parity"] = {"checked": fx is not None, "tests": checked, "differing": bad,
                     "against": "GMP (tests/golden/prp_n2e10_b2e19.json): RES64 of base^(N-1) mod N"}
    return out


def bytes_alg(n):
    """Algorithmic HBM bytes per squaring of one candidate: 24 n R (R = 1 if the
    state fits one CU, else 2 exchanges)."""
    return 24 * n * (1 if n <= (1 << 13) else 2)


BATCH_FIXTURE = os.path.join(ROOT, "tests", "golden", "batched_n2e14_x256.json")


def batch_parity(ct, t, bs, done, i0):
    """After the timed run (`done` squarings from x = 1): continue untimed to
    the next GMP checkpoint of tests/golden/batched_n2e14_x256.json and compare
    RES64 and x mod three 31-bit primes of every lane's canonical residue."""
    if not os.path.exists(BATCH_FIXTURE):
        return {"checked": False, "reason": "fixture missing"}
    fx = json.load(open(BATCH_FIXTURE))
    if [L["b"] for L in fx["lanes"]] != list(bs):
        return {"checked": False, "reason": "fixture drawn for other b"}
    later = [tc for tc in fx["t"] if tc >= done]
    if not later:
        return {"checked": False, "reason": f"no GMP checkpoint at or after {done} squarings"}
    tc = later[0]
    if tc > done:
        t.run(i0 - done, i0 - tc + 1)
    r64, mods = ct.residue_checks(t.get_int(0), bs, fx["q"])
    k = fx["t"].index(tc)
    bad = [l for l, L in enumerate(fx["lanes"])
           if "%016X" % r64[l] != L["aligned_res64"][k] or mods[l] != L["aligned_mod_q"][k]]
    return {"checked": True, "lanes": len(bs), "squarings": tc, "timed_squarings": done,
            "against": "GMP (tests/golden/batched_n2e14_x256.json): RES64 and x mod 3 primes",
            "lanes_differing": len(bad), "bit_exact": not bad}


def run_batched(ct, steps, warmup, csize=0):
    """`warmup` and `steps` count squarings here (the caller converts bench
    steps of SQ_PER_STEP squarings).  256 lanes, each from its own top
    exponent bit: lane l's exponent is shifted left so that its top bit sits
    at the common index i0, so every
    lane carries a non-trivial residue from the first squaring (in a lockstep
    batch from a common i0, lanes with a shorter exponent square x = 1 for up
    to thousands of steps).  After `steps` timed squarings lane l holds
    2^(top `steps` bits of E_l) mod N_l."""
    from ctwin_amd.exponent import exponent
    ln, n, lanes = 14, 1 << 14, 256
    rng = random.Random(SEED)
    bs = [rng.randrange(1 << 18, 1 << 19) for _ in range(lanes)]
    E = [exponent(b, n) for b in bs]
    top = max(e.bit_length() for e in E)
    i0 = top - 1
    aligned = [e << (top - e.bit_length()) for e in E]
    # configuration chosen by the autotuner (timed before this run, outside the timed region)
    t = ct.Transform(ln, lanes, ct.KIND_AUTO, csize=csize)
    t.set_b(bs); t.set_exponents(aligned); t.set(1)
    t.run(i0, i0 - warmup + 1)                       # warm-up, then the state restarts from x = 1
    t.set(1)
    t.timer_start()
    t.run(i0, i0 - steps + 1)
    ms = t.timer_stop()
    kind, tile = t.kind, t.tuning()
    parity = batch_parity(ct, t, bs, steps, i0)
    t.close()
    return lanes * steps / (ms / 1e3), ms, i0, bs, kind, tile, parity


def prp_work(i0):
    """Squaring-equivalents of one PRP test with Gerbicz-Li (cyclo.h:252-377)
    whose exponent's top bit is i0: the loop's i0 + 1 squarings, L = 2^(1 +
    floor(log2 i0)/2) products d(t) x (one product costs about two squarings:
    two forward transforms), the check's B + 11 squarings (B = ceil(i0 / L),
    the top bit of the B-bit chunk sum is below B + log2 L + 1) and its two
    products."""
    L = 2 << ((i0.bit_length() - 1) // 2)
    B = (i0 - 1) // L + 1
    return (i0 + 1) + 2 * L + (B + 11) + 4


LARGE_FIXTURE = os.path.join(ROOT, "tests", "golden", "large_n2e21_b73764.jsonl")


def large_parity(ct, t, b, ln, done):
    """After `done` squarings from x = 1 at the exponent's top bit: RES64 and
    x mod three 31-bit primes of the canonical residue against GMP's
    checkpoint at t = done (tests/golden/large_n2e21_b73764.jsonl)."""
    if not os.path.exists(LARGE_FIXTURE):
        return {"checked": False, "reason": "fixture missing"}
    lines = [json.loads(l) for l in open(LARGE_FIXTURE) if l.strip()]
    head, checks = lines[0], {c["t"]: c for c in lines[1:]}
    if head["b"] != b or head["log2_n"] != ln or done not in checks:
        return {"checked": False, "reason": f"no GMP checkpoint for b={b}, n=2^{ln}, t={done}"}
    r64, mods = ct.residue_checks(t.get_int(0), [b], head["q"])
    want = checks[done]
    ok = "%016X" % r64[0] == want["res64"] and mods[0] == want["mod_q"]
    return {"checked": True, "squarings": done, "bit_exact": ok,
            "against": "GMP (tests/golden/large_n2e21_b73764.jsonl): RES64 and x mod 3 primes"}


def run_plain(ct, ln, b, steps, warmup, graphs=True, config=None, check=False):
    """Plain multipass engine, one candidate: `warmup` untimed squarings, then
    `steps` timed ones.  The warm-up also runs one run of the timed length, so
    that the captured graph of that length exists (a long PRP test replays it
    millions of times).  config = (tile, split, fused) instead of the autotuner.
    check: the timed run restarts from x = 1 at the exponent's top bit (a
    squaring costs the same whatever the digits) and its result is compared
    with GMP (large_parity)."""
    from ctwin_amd.exponent import window_value
    n = 1 << ln
    Ew, i0 = window_value(b, n, 2 * steps + warmup + 8)
    if config is None:
        t = ct.Transform(ln, 1, ct.KIND_PLAIN, csize=0)     # multipass, tile and pass split from the autotuner
    else:
        t = ct.Transform(ln, 1, ct.KIND_PLAIN)
        t.set_column_tile(config[0]); t.set_column_split(config[1]); t.set_fused(config[2])
    t.set_graphs(graphs)
    t.set_b([b]); t.set_exponents([Ew]); t.set(1)
    t.run(i0, i0 - warmup + 1)
    t.run(i0 - warmup, i0 - warmup - steps + 1)
    i = i0 - warmup - steps
    if check:
        t.set(1)
        i = i0
    t.timer_start()
    t.run(i, i - steps + 1)
    ms = t.timer_stop()
    w, split, fused, _, row_form = t.tuning()
    t.read_state(0)                 # the read checks the fused pass's carry exchanges
    cfg = {"column_tile": w, "split_pass_a": split, "fused_pass_ca": fused, "row_form": row_form}
    if check:
        cfg["parity"] = large_parity(ct, t, b, ln, steps)
    t.close()
    return steps / (ms / 1e3), ms, cfg


def run_sieve_window(ct):
    """Window sieve: m = 12, b in [2^20, 2^20 + 2^24), all p < 2^32, both sides.
    One untimed call first (it creates the device buffers and loads the
    kernels), then the timed call; survivors against the reference program's
    list, the work of the reference algorithm on this window
    (tests/golden/sieve_work_m12_window_2e24.json) per second, and the VALU roofline of
    the N- root kernel from its live launch times."""
    import hashlib
    import numpy as np
    m, b_min, b_max = 12, 1 << 20, (1 << 20) + (1 << 24) - 1
    t0 = time.perf_counter()
    ct.sieve_window(m, b_min, b_max, 1 << 32, 1 << 32)
    first = time.perf_counter() - t0
    t0 = time.perf_counter()
    st = ct.sieve_window(m, b_min, b_max, 1 << 32, 1 << 32)
    dt = time.perf_counter() - t0
    timing = ct.sieve_last_timing()
    surv = ct.survivors(st, b_min)
    out = {"seconds": dt, "first_call_seconds": first, "survivors": int(len(surv)), "b_window": 1 << 24,
           "p_max": 1 << 32, "b_per_s": (1 << 24) / dt, "device": timing}
    fx = os.path.join(ROOT, "tests", "golden", "sieve_m12_window_2e24.json")
    if os.path.exists(fx):
        want = json.load(open(fx))
        text = "".join(f"{int(b)}\n" for b in surv)
        out["bit_exact_vs_reference"] = hashlib.sha256(text.encode()).hexdigest() == want["sha256"]
    wk = os.path.join(ROOT, "tests", "golden", "sieve_work_m12_window_2e24.json")
    if os.path.exists(wk):
        work = json.load(open(wk))
        total = work["montgomery_products"] + work["verify_products"]
        out["reference_algorithm_work"] = {
            "montgomery_products": work["montgomery_products"], "verify_products": work["verify_products"],
            "products_per_s": total / dt,
            "note": "products the reference's CPU algorithm (ctsieve.cpp check_pos/check_neg/check_root) "
                    "spends on this window, counted by its restatement; this build computes the same "
                    "survivors with a segmented sieve and direct roots instead"}
    launches = max(timing["roots_neg_launches"], 1)
    per_launch_s = timing["roots_neg_ms"] / 1e3 / launches
    achieved = SIEVE_ROOTS_NEG_SLOTS / launches / per_launch_s / 1e9 if per_launch_s > 0 else None
    out["roofline"] = {"bound": "valu", "kernel": "k_roots_neg", "achieved": achieved, "peak": VALU_SLOT_PEAK_G,
                       "unit": "G VALU issue slots/s", "frac": achieved / VALU_SLOT_PEAK_G if achieved else None,
                       "traffic": TRAFFIC_SIEVE_ROOTS_NEG_LAUNCH, "launch_ms": per_launch_s * 1e3,
                       "note": "slots per launch (counted VALU instructions minus dual-issue quad-cycles, 26 "
                               "launches on this window) over the launch time measured with HIP events; the "
                               "kernel is integer-VALU bound, its HBM traffic is the prime list and the strikes"}
    return out


def sieve_pool(ct, m, bits, cache):
    """Sieve survivors (both twin members, p < 2^26) of the window
    [2^(bits-1), 2^(bits-1) + 2^16) for n = 2^m digits: the b values the served
    stream draws from."""
    key = (m, bits)
    if key not in cache:
        lo = 1 << (bits - 1)
        hi = lo + min(1 << 16, lo) - 1
        st = ct.sieve_window(m, lo, hi, 1 << 26, 1 << 26)
        cache[key] = [int(b) for b in ct.survivors(st, lo)]
    return cache[key]


def run_served(ct, requests=72, steps=1000, rate=400.0, budget=16 << 30, graphs=True, twin=False):
    """Served stream: Poisson arrivals, n = 2^k with k uniform in 12..20, b of 16..22
    bits taken from sieve survivors, 1000 squarings per request, admission by an
    HBM budget.  twin: each arrival is a candidate pair, its N+ and N- tests
    submitted together (requests / 2 pairs)."""
    import numpy as np
    rng = random.Random(SEED + 5)
    pools = {}
    t_arr, reqs, t = [], [], 0.0
    signs = (1, -1) if twin else (1,)
    for i in range(requests // len(signs)):
        t += rng.expovariate(rate / len(signs))
        bits = rng.randrange(16, 23)
        k = rng.randrange(12, 21)
        b = rng.choice(sieve_pool(ct, k, bits, pools))
        for sg in signs:
            reqs.append((len(reqs), k, b, steps, sg))
            t_arr.append(t)
    s = ct.Server(hbm_budget=budget, slots=16, chunk=125, max_steps=steps)
    s.set_graphs(graphs)
    # warm-up: one short request per size class creates the class engines
    # (autotuned kind and tile, tables, buffers) as a long-running server would
    # already hold them
    for k in range(12, 21):
        for sg in signs:
            s.submit(10 ** 6 + 2 * k + (sg < 0), k, 65537, 8, sg)
    warm = 0
    while warm < 9 * len(signs):
        s.step()
        warm += len(s.poll())
    sq0 = s.stats()["squarings"]
    lat = {}
    start = time.perf_counter()
    nxt = 0
    while len(lat) < requests:
        now = time.perf_counter() - start
        while nxt < requests and t_arr[nxt] <= now:
            s.submit(*reqs[nxt]); nxt += 1
        s.step()
        for rid, r64, h, l in s.poll():
            lat[rid] = l
        if nxt < requests and s.stats()["active"] == 0 and s.stats()["queued"] == 0:
            time.sleep(max(0.0, t_arr[nxt] - (time.perf_counter() - start)))
    wall = time.perf_counter() - start
    st = s.stats()
    s.close()
    L = np.array(sorted(lat.values()))
    per_class = {}
    for k in range(12, 21):
        for sg in signs:
            v = np.array([lat[r[0]] for r in reqs if r[1] == k and r[4] == sg])
            if len(v):
                name = f"2^{k}" + (("+" if sg > 0 else "-") if twin else "")
                per_class[name] = {"requests": int(len(v)), "p50_ms": 1e3 * float(np.percentile(v, 50)),
                                   "p99_ms": 1e3 * float(np.percentile(v, 99))}
    return {"requests": len(reqs), "twin_pairs": twin, "squarings_per_request": steps, "arrival_rate_per_s": rate,
            "wall_s": wall, "squarings_per_s": (st["squarings"] - sq0) / wall,
            "requests_per_s": requests / wall,
            "latency_p50_ms": 1e3 * float(np.percentile(L, 50)), "latency_p99_ms": 1e3 * float(np.percentile(L, 99)),
            "latency_by_class": per_class,
            "hbm_budget_bytes": budget, "log2_n_range": [12, 20], "b_bits_range": [16, 22],
            "b_source": "GPU sieve survivors (both members, p < 2^26) of 2^16-wide windows"}


def _cpu_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = min(avail, int(omp)) if omp and omp.isdigit() else avail
    return model, os.cpu_count(), avail, max(1, threads)


def _cpu_rate(C, eng, x, bits, words, i0, threads, lanes, budget_s, chunk):
    """Squarings/s of the CPU comparator: runs of `chunk` PRP-loop steps until
    budget_s has elapsed or the exponent window (bits i0 .. 0) is used up
    (at least one run)."""
    done, t0 = 0, time.perf_counter()
    while True:
        eng.steps(x, bits, words, i0 - done, chunk, threads)
        done += chunk
        dt = time.perf_counter() - t0
        if dt >= budget_s or done + chunk > i0 + 1:
            return lanes * done / dt, lanes * done, dt


def cpu_baseline(ct, budget_s=5.0):
    """CPU baselines on the GPU host (the cpu_baseline leg):
    - the C++ comparator (oracle/cpu_port: the same two-prime NTT, Garner and
      chunk carry as the engine, OpenMP) on bounded samples of the batched
      workload (n = 2^14, lanes of the bench's batch) and of the single large
      candidate (n = 2^21), one thread and every thread this process may use;
    - the reference's own CPU window sieve (oracle/_ref/ctsieve_ref, compiled
      from sieve/CPU/ctsieve.cpp) on the m = 12 window's b with p < 2^28, beside
      the GPU sieve on the same sample (same survivors required)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import hashlib
    import subprocess
    import tempfile
    import numpy as np
    import cpu_port as C
    from ctwin_amd.exponent import exponent, top_bits
    model, nproc, avail, threads = _cpu_info()
    rng = random.Random(SEED)
    bs = [rng.randrange(1 << 18, 1 << 19) for _ in range(256)]
    n = 1 << 14
    out = {"unit": "squarings/s", "kind": "port", "cpu_model": model, "nproc": nproc, "cpus_available": avail}
    rates = {}
    for th in (1, threads):
        lanes = 2 * th
        E = [exponent(b, n) for b in bs[:lanes]]
        top = max(e.bit_length() for e in E)
        bits, words = C.pack_bits([e << (top - e.bit_length()) for e in E], lanes)
        eng = C.CpuEngine(14, bs[:lanes])
        x = np.zeros((lanes, n), np.int32); x[:, 0] = 1
        r, cnt, dt = _cpu_rate(C, eng, x, bits, words, top - 1, th, lanes, budget_s, 4)
        rates[th] = r
        out.setdefault("n2e14", {})[f"threads_{th}"] = {"squarings_per_s": r, "sample": f"{cnt} squarings ({lanes} lanes) in {dt:.1f} s"}
        eng.close()
    big_b = 65536 + 2 * random.Random(SEED).randrange(1 << 14)
    w, sh = top_bits(big_b, 1 << 21, 4096)
    bits, words = C.pack_bits([w], 1)
    for th in (1, threads):
        eng = C.CpuEngine(21, [big_b])
        x = np.zeros((1, 1 << 21), np.int32); x[0, 0] = 1
        r, cnt, dt = _cpu_rate(C, eng, x, bits, words, 4095, th, 1, budget_s, 1)
        out.setdefault("n2e21", {})[f"threads_{th}"] = {"squarings_per_s": r, "sample": f"{cnt} squarings in {dt:.1f} s"}
        eng.close()
    out["value"] = rates[threads]
    out["cores"] = threads
    out["sample"] = (f"C++ comparator, n=2^14 batch lanes, {threads} OpenMP threads over lanes, "
                     f"{budget_s:.0f} s per point")
    # the reference's CPU sieve on a bounded sample, and the GPU sieve beside it
    ref = os.path.join(ROOT, "oracle", "_ref", "ctsieve_ref")
    m, b_min, b_max, pm = 12, 1 << 20, (1 << 20) + (1 << 24) - 1, 1 << 28
    sv = {"kind": "reference", "sample": f"m=12, b in [2^20, 2^20 + 2^24), p < 2^28 both sides, 1 thread "
                                         "(sieve/CPU/ctsieve.cpp, its own build)"}
    if os.path.exists(ref):
        with tempfile.TemporaryDirectory() as d:
            f = os.path.join(d, "surv.txt")
            t0 = time.perf_counter()
            rc = subprocess.run([ref, str(m), str(b_min), str(b_max), str(pm), str(pm), f],
                                stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL).returncode
            dt = time.perf_counter() - t0
            ref_text = open(f).read() if rc == 0 else None
        t0 = time.perf_counter()
        st = ct.sieve_window(m, b_min, b_max, pm, pm)
        gdt = time.perf_counter() - t0
        gpu_text = "".join(f"{int(b)}\n" for b in ct.survivors(st, b_min))
        sv.update({"reference_seconds": dt, "gpu_seconds": gdt, "reference_rc": rc,
                   "survivors": gpu_text.count("\n"),
                   "gpu_equals_reference": ref_text is not None and
                   hashlib.sha256(ref_text.encode()).digest() == hashlib.sha256(gpu_text.encode()).digest()})
    else:
        sv["skipped"] = "oracle/_ref/ctsieve_ref not built"
    out["sieve"] = sv
    out["sieve_m12_window_reference_s_build_container"] = {
        "seconds": 177.5, "where": "the build container (Xeon 2.7 GHz, 1 thread), not the GPU host",
        "sample": "the full m = 12 window of 2^24 b (p < 2^32 both sides)"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="timed steps of the batched workload, each SQ_PER_STEP squarings of all 256 candidates")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--plain-steps", type=int, default=300)
    ap.add_argument("--big-steps", type=int, default=10000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graphs", action="store_true",
                    help="multipass engines launch every kernel instead of replaying captured HIP graphs "
                         "(same results; used for kernel traces)")
    ap.add_argument("--sieve-only", action="store_true",
                    help="run only the m = 12 window sieve (2^24 b, p < 2^32) and print its JSON object (for profiling the sieve)")
    ap.add_argument("--profile", action="store_true",
                    help="short run for hardware counters, without graphs or autotuning: 50 batched steps "
                         "in one resident launch, 20 plain squarings at n=2^14 (column tile 16) and n=2^21 "
                         "(tile 32), fused column pass with the primes in workgroup halves, and 3 copies of "
                         "an 8 MiB register (calibration of FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the candidates shard one queue per GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import ctwin_amd as ct
    if args.no_graphs or args.profile:
        ct.set_graphs_default(False)          # no graph anywhere in the process, the autotuner's included

    if args.sieve_only:
        print(json.dumps({"sieve_m12_window": run_sieve_window(ct)}), flush=True)
        return

    if args.profile:
        run_batched(ct, 50, 1, csize=16)[:0]
        b0 = random.Random(SEED).randrange(1 << 18, 1 << 19)
        run_plain(ct, 14, b0, 20, 1, graphs=False, config=(16, True, 1))
        run_plain(ct, 21, 65536 + 2 * random.Random(SEED).randrange(1 << 14), 20, 1, graphs=False, config=(32, True, 1))
        t = ct.Transform(21, 1, ct.KIND_PLAIN)
        for _ in range(3):
            t.copy(1, 0)
        t.timer_start(); t.timer_stop()
        print("[bench] profile run done (copy kernel moves 8388608 bytes each way)", flush=True)
        return

    batch_sq = args.steps * SQ_PER_STEP
    sq_s, ms, i0, bs, kind, tuning, parity = run_batched(ct, batch_sq, max(1, args.warmup * SQ_PER_STEP))
    print(f"[bench] batched n=2^14 x256: {sq_s:.4g} squarings/s ({ms:.1f} ms for {args.steps} steps of "
          f"{SQ_PER_STEP} squarings); "
          f"parity {parity}", flush=True)
    plain_sq_s, plain_ms, plain_cfg = run_plain(ct, 14, bs[0], args.plain_steps, 5, graphs=not args.no_graphs)
    print(f"[bench] plain n=2^14 x1: {plain_sq_s:.4g} squarings/s", flush=True)
    big_b = 65536 + 2 * random.Random(SEED).randrange(1 << 14)
    big_sq_s, big_ms, big_cfg = run_plain(ct, 21, big_b, args.big_steps, 5, graphs=not args.no_graphs, check=True)
    big_parity = big_cfg.pop("parity")
    print(f"[bench] plain n=2^21 x1: {big_sq_s:.4g} squarings/s; parity {big_parity}", flush=True)

    sieve = run_sieve_window(ct)
    print(f"[bench] sieve m = 12, 2^24 b: {sieve['survivors']} survivors in {sieve['seconds']:.3f} s", flush=True)
    served = run_served(ct, graphs=not args.no_graphs)
    print(f"[bench] served: {served['squarings_per_s']:.4g} squarings/s, p50 {served['latency_p50_ms']:.1f} ms, "
          f"p99 {served['latency_p99_ms']:.1f} ms", flush=True)
    served_light = run_served(ct, requests=48, rate=40.0, graphs=not args.no_graphs)
    print(f"[bench] served below saturation: {served_light['squarings_per_s']:.4g} squarings/s, "
          f"p50 {served_light['latency_p50_ms']:.1f} ms, p99 {served_light['latency_p99_ms']:.1f} ms", flush=True)
    served_twin = run_served(ct, twin=True, graphs=not args.no_graphs)
    print(f"[bench] served twin pairs (N+ and N-): {served_twin['squarings_per_s']:.4g} squarings/s, "
          f"p50 {served_twin['latency_p50_ms']:.1f} ms, p99 {served_twin['latency_p99_ms']:.1f} ms", flush=True)

    prp10 = run_prp_n2e10(ct)
    print(f"[bench] full PRP n = 2^10: {prp10['plus_base3_batched1024']['candidates_per_hour']:.4g} N+ tests/hour (1024 lanes), "
          f"{prp10['plus_base3_batched64']['candidates_per_hour']:.4g} (64 lanes), parity {prp10['parity']}", flush=True)
    n = 1 << 14
    per_cand = prp_work(i0)
    valu_achieved = mul32_alg(n, 2) * sq_s / 1e12
    valu_issue = VALU_SLOTS_PER_SQUARING / (VALU_SLOT_PEAK_G * 1e9) / (ms / 1e3 / batch_sq)
    big_n = 1 << 21
    hbm_achieved = bytes_alg(big_n) * big_sq_s / 1e9
    out = {
        "metric": "Fermat squarings/s (batched, n=2^14 digits, bit-exact)",
        "value": sq_s,
        "unit": "squarings/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded b)",
        "config": {"workload": "batched: 256 candidates b^(2^14) - b^(2^13) + 1, b in [2^18, 2^19), "
                               "base-2 PRP loop from each lane's own top exponent bit (every residue "
                               "non-trivial), one persistent launch of the autotuned engine",
                   "log2_n": 14, "candidates": 256, "seed": SEED, "squarings_per_step": SQ_PER_STEP,
                   "squarings_timed": batch_sq,
                   "engine": "resident" if kind == ct.KIND_RESIDENT else "multipass",
                   "autotuner_ms_per_step": dict(zip(TUNE_NAMES, tuning[3]))},
        "candidates_per_hour": 3600.0 * sq_s / per_cand,
        "parity": parity,
        "plain": {"squarings_per_s": plain_sq_s, "candidates_per_hour": 3600.0 * plain_sq_s / per_cand,
                  "engine": "plain multipass, one candidate per launch sequence", "tuned": plain_cfg},
        "batched_vs_plain": sq_s / plain_sq_s,
        "single_large": {"log2_n": 21, "b": big_b, "squarings_per_s": big_sq_s, "ms_per_squaring": big_ms / args.big_steps,
                         "tuned": big_cfg, "parity": big_parity,
                         "roofline": {"bound": "hbm", "achieved": hbm_achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                      "frac": hbm_achieved / HBM_PEAK_GBS, "traffic": TRAFFIC_N2E21_SQUARING,
                                      "bytes_alg_per_squaring": bytes_alg(big_n),
                                      "valu_issue_frac": VALU_SLOTS_N2E21_SQUARING / (VALU_SLOT_PEAK_G * 1e9)
                                                         / (big_ms / 1e3 / args.big_steps),
                                      "note": "bound by neither: HBM traffic 1.2x the algorithmic bytes at about a "
                                              "third of 8 TB/s, VALU issue (row pass + fused column pass slots "
                                              "over the measured squaring time) a little above half"}},
        "sieve_m12_window": sieve,
        "served": served,
        "served_below_saturation": served_light,
        "served_twin_pairs": served_twin,
        "prp_n2e10": prp10,
        "roofline": {"bound": "valu", "achieved": valu_achieved, "peak": VALU_PEAK_T, "unit": "Tmul32/s",
                     "frac": valu_achieved / VALU_PEAK_T, "traffic": TRAFFIC_RESIDENT_LAUNCH,
                     "kernel": "k_res_steps<14, 2>", "primes": 2, "mul32_alg_per_squaring": mul32_alg(n, 2),
                     "valu_issue_frac": valu_issue,
                     "note": "achieved prices the two-prime work the kernel does (6 n (log2 n + 3) mul32 per "
                             "squaring); valu_issue_frac = counted VALU issue slots per step (instructions "
                             "minus dual-issue quad-cycles) over 614.4 G slots/s x the measured step time; "
                             "traffic = HBM bytes per launch from FETCH_SIZE/WRITE_SIZE (state and tables "
                             "read once per launch)"},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(ct)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
