"""Gerbicz-Li failure and quit paths on the HIP engines.

- A PRP test is split at its loop (ct_prp_begin / ct_prp_finish), one digit of
  register 0 of one lane is altered from the host between two pieces of the
  loop, and the verification must fail for that lane only (CT_ERR_GERBICZ,
  gl_ok = 0), as the reference's fault hook makes it fail (cyclo.h:298).  The
  untouched run of the same job passes and its RES64 equals GMP's.  Resident
  kernels, the multipass engine and its fused column pass.
- The batch driver is stopped by ct_cyclo_quit() from another thread while it
  runs (the reference's SIGINT flag, cyclo.h:292-296, main.cpp:33-36), resumes
  from its checkpoint and writes the same result_c.txt as an uninterrupted run."""
import os
import random
import threading
import time

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
ct = pytest.importorskip("ctwin_amd")


def _run_job(t, bs, corrupt_lane=None):
    job = ct.PrpJob(t, bs)
    i0, B = job.i0, job.gl_block
    mid = i0 - 3 * B - B // 2                      # inside a Gerbicz-Li block
    t.set(1)
    t.copy(1, 0)
    t.run(i0, mid, B)
    if corrupt_lane is not None:
        x = t.read_state(0)
        j = 5
        x[corrupt_lane, j] = x[corrupt_lane, j] + 1 if x[corrupt_lane, j] < bs[corrupt_lane] - 1 else x[corrupt_lane, j] - 1
        t.write_state(0, x)
    t.run(mid - 1, 0, B)
    out = job.finish()
    job.close()
    return out


@pytest.mark.parametrize("ln,kind,fused", [(10, 1, 0), (10, 2, 0), (13, 2, 1)])
def test_gerbicz_detects_altered_register(ln, kind, fused):
    """The untouched run's RES64 equals GMP's (n = 2^10: 20k-bit powm) or, for
    the fused column pass at n = 2^13, the resident kernels' RES64 of the same
    job (those kernels are GMP-pinned at this size elsewhere)."""
    n, lanes = 1 << ln, 4
    rng = random.Random(300 + ln + fused)
    bs = [rng.randrange(1 << 16, 1 << 19) for _ in range(lanes)]
    print(f"\n[case] gerbicz ln={ln} kind={kind} fused={fused} bs={bs}", flush=True)
    t = ct.Transform(ln, lanes, kind)
    if kind == ct.KIND_PLAIN:
        t.set_fused(fused)
    prm, r64, gl, _, rc = _run_job(t, bs)
    assert rc == 0 and gl.all()
    if ln <= 10:
        for l, b in enumerate(bs):
            N = O.modulus(n, b)
            assert int(r64[l]) == O.powm(2, N - 1, N) % (1 << 64), f"lane {l} RES64 vs GMP"
    else:
        ref = ct.Transform(ln, lanes, ct.KIND_RESIDENT)
        _, r64_res, gl_res, _, rc_res = _run_job(ref, bs)
        ref.close()
        assert rc_res == 0 and gl_res.all() and list(r64_res) == list(r64)
    print(f"[gerbicz] untouched run passes, RES64 {[ct.res64_string(v) for v in r64]}", flush=True)
    prm2, r64b, gl2, _, rc2 = _run_job(t, bs, corrupt_lane=2)
    assert rc2 == -7, f"status {rc2}, expected CT_ERR_GERBICZ"
    assert list(gl2) == [True, True, False, True]
    assert r64b[2] != r64[2] and all(r64b[l] == r64[l] for l in (0, 1, 3))
    print("[gerbicz] altered lane 2 detected, the others pass", flush=True)
    t.close()


def test_quit_from_another_thread_then_resume(tmp_path):
    ln, lanes = 12, 8
    rng = random.Random(4242)
    bs = [rng.randrange(1 << 17, 1 << 18) for _ in range(24)]      # three batches
    cand = tmp_path / "cand.txt"
    cand.write_text("".join(f"{b}\n" for b in bs))
    print(f"\n[case] quit/resume ln={ln} lanes={lanes} candidates={len(bs)}", flush=True)
    a = tmp_path / "a"; a.mkdir()
    rc, st = ct.cyclo_check_file(ln, str(cand), str(a), lanes=lanes)
    assert rc == 0 and st["candidates"] == len(bs)
    b = tmp_path / "b"; b.mkdir()
    out = []
    th = threading.Thread(target=lambda: out.append(
        ct.cyclo_check_file(ln, str(cand), str(b), lanes=lanes, checkpoint_s=0.02)))
    th.start()
    ctx = b / "cyclo.ctx"
    t0 = time.time()
    while not ctx.exists() and th.is_alive() and time.time() - t0 < 60:
        time.sleep(0.005)
    assert ctx.exists(), "the driver wrote no checkpoint while running"
    ct.cyclo_quit()
    th.join(120)
    assert not th.is_alive()
    rc, st = out[0]
    assert rc == ct.INTERRUPTED and ctx.exists()
    print(f"[quit] stopped after {st['squarings']} squarings", flush=True)
    rc, st2 = ct.cyclo_check_file(ln, str(cand), str(b), lanes=lanes)
    assert rc == 0 and st2["resumed"]
    assert (b / "result_c.txt").read_text() == (a / "result_c.txt").read_text()
    assert not ctx.exists()
