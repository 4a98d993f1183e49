"""ctwin_amd: MI355X-native engines for galloty/ctwin's data-parallel paths.

Python mirror of the reference's host interfaces over the C ABI of
include/ctwin.h (libctwin_amd.so, built in-tree by ctwin_amd/build.py):

* Transform   -- cyclo/src/transform.h + engine.h: set_b, set, copy,
                 square_dup, mul, get_int, is_prime, context state.
* prp_check   -- cyclo/src/cyclo.h:252-411, base-2 Fermat PRP with Gerbicz-Li.
* sieve_window -- sieve/CPU/ctsieve.cpp Sieve::check_pos/check_neg.

The library is required: importing the engines without it raises, there is
no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libctwin_amd.so")

KIND_AUTO, KIND_RESIDENT, KIND_PLAIN = 0, 1, 2
TUNE_TIMINGS = 13                    # CT_TUNE_TIMINGS

STATUS = {0: "CT_OK", -1: "CT_ERR_ARG", -2: "CT_ERR_HIP", -3: "CT_ERR_NOMEM", -4: "CT_ERR_UNSUPPORTED",
          -5: "CT_ERR_GETINT", -6: "CT_ERR_COMPUTE", -7: "CT_ERR_GERBICZ"}

EXPORTS = ["ct_engine_create", "ct_engine_destroy", "ct_engine_set_b", "ct_engine_set", "ct_engine_copy",
           "ct_engine_square_dup", "ct_engine_mul", "ct_engine_mul_regs", "ct_engine_get_int",
           "ct_engine_is_prime", "ct_engine_read_state", "ct_engine_write_state", "ct_engine_set_exponents",
           "ct_engine_run", "ct_prp_check", "ct_engine_timer_start", "ct_engine_timer_stop", "ct_engine_info",
           "ct_sieve_window", "ct_last_error", "ct_engine_set_primes", "ct_engine_primes",
           "ct_server_create", "ct_server_destroy", "ct_server_submit", "ct_server_submit_ex", "ct_server_step", "ct_server_poll",
           "ct_server_stats", "ct_sieve_factors", "ct_engine_moduli",
           "ct_engine_set_graphs", "ct_engine_set_column_tile", "ct_engine_tuning", "ct_crc32", "ct_sieve_save",
           "ct_sieve_load", "ct_context_save", "ct_context_load", "ct_cyclo_check_file", "ct_cyclo_quit", "ct_server_set_graphs",
           "ct_engine_create_ex", "ct_engine_set_base", "ct_engine_set_column_split",
           "ct_engine_set_fused", "ct_prp_begin", "ct_prp_finish", "ct_prp_job_free",
           "ct_sieve_factors_noalloc", "ct_engine_set_row_form", "ct_engine_set_prime_halves", "ct_set_graphs_default",
           "ct_fsieve_create", "ct_fsieve_blocks", "ct_fsieve_set_bmax", "ct_fsieve_run", "ct_fsieve_destroy",
           "ct_fsieve_check", "ct_fsieve_quit", "ct_sieve_last_timing", "ct_merge_factors"]

INTERRUPTED = 1


class CycloStats(ctypes.Structure):
    _fields_ = [("candidates", ctypes.c_uint64), ("prps", ctypes.c_uint64), ("squarings", ctypes.c_uint64),
                ("elapsed", ctypes.c_double), ("resumed", ctypes.c_int)]


class FsieveStats(ctypes.Structure):
    _fields_ = [("blocks_done", ctypes.c_uint64), ("blocks_total", ctypes.c_uint64), ("factors", ctypes.c_uint64),
                ("elapsed", ctypes.c_double), ("resumed", ctypes.c_int), ("interrupted", ctypes.c_int)]


class MergeStats(ctypes.Structure):
    _fields_ = [("candidates", ctypes.c_uint64), ("pairs", ctypes.c_uint64), ("in_window", ctypes.c_uint64),
                ("factors", ctypes.c_uint64), ("bad", ctypes.c_uint64), ("out_of_range", ctypes.c_uint64),
                ("remaining", ctypes.c_uint64)]


class SieveTiming(ctypes.Structure):
    _fields_ = [("roots_neg_ms", ctypes.c_float), ("roots_neg_launches", ctypes.c_uint32), ("device_ms", ctypes.c_float),
                ("kernels_ms", ctypes.c_float)]


class CtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def lib():
    """Loads libctwin_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m ctwin_amd.build` (no CPU fallback exists)")
        L = ctypes.CDLL(LIB_PATH)
        c = ctypes
        vp, u32, i32, u64 = c.c_void_p, c.c_uint32, c.c_int, c.c_uint64
        u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
        sig = {
            "ct_engine_create": [i32, u32, u32, u32, i32, c.POINTER(vp)],
            "ct_engine_destroy": [vp],
            "ct_engine_set_b": [vp, u32p],
            "ct_engine_set": [vp, u32],
            "ct_engine_copy": [vp, u32, u32],
            "ct_engine_square_dup": [vp, c.c_void_p],
            "ct_engine_mul": [vp],
            "ct_engine_mul_regs": [vp, u32, u32, u32],
            "ct_engine_get_int": [vp, u32, u32p],
            "ct_engine_is_prime": [vp, u32p, ip, u64p],
            "ct_engine_read_state": [vp, u32, i32p],
            "ct_engine_write_state": [vp, u32, i32p],
            "ct_engine_set_exponents": [vp, u32p, u32],
            "ct_engine_run": [vp, i32, i32, i32],
            "ct_prp_check": [vp, u32p, ip, u64p, ip, c.c_void_p],
            "ct_engine_timer_start": [vp],
            "ct_engine_timer_stop": [vp, c.POINTER(c.c_float)],
            "ct_engine_info": [vp, c.POINTER(u32), c.POINTER(u32), c.POINTER(i32)],
            "ct_sieve_window": [i32, i32, u64, u64, u64, u64, u64, u64, i32, u8p],
            "ct_sieve_factors": [i32, i32, u64, u64, i32, u64p, u64, c.POINTER(u64)],
            "ct_engine_set_primes": [vp, i32],
            "ct_server_create": [i32, u64, u32, u32, u32, c.POINTER(vp)],
            "ct_server_destroy": [vp],
            "ct_server_submit": [vp, u64, u32, u32, u32],
            "ct_server_submit_ex": [vp, u64, u32, u32, u32, i32],
            "ct_server_step": [vp],
            "ct_server_poll": [vp, u64p, u64p, u64p, np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS"), u32,
                               c.POINTER(u32)],
            "ct_server_stats": [vp, c.POINTER(u32), c.POINTER(u32), c.POINTER(u64), c.POINTER(u64)],
            "ct_engine_primes": [vp, c.POINTER(i32)],
            "ct_engine_moduli": [u32p],
            "ct_engine_create_ex": [i32, u32, u32, u32, i32, i32, c.POINTER(vp)],
            "ct_engine_set_base": [vp, u32],
            "ct_server_set_graphs": [vp, i32],
            "ct_engine_set_graphs": [vp, i32],
            "ct_engine_set_column_tile": [vp, u32],
            "ct_sieve_save": [c.c_char_p, i32, u64, u64, u64, u64, u8p, i32],
            "ct_sieve_load": [c.c_char_p, i32, u64, u64, c.POINTER(u64), c.POINTER(u64), u8p],
            "ct_context_save": [vp, c.c_char_p, i32, i32, c.c_double],
            "ct_context_load": [vp, c.c_char_p, c.POINTER(i32), c.POINTER(i32), c.POINTER(c.c_double)],
            "ct_cyclo_check_file": [i32, u32, c.c_char_p, c.c_char_p, u32, c.c_double, u64, c.POINTER(CycloStats)],
            "ct_engine_tuning": [vp, c.POINTER(u32), c.POINTER(i32), c.POINTER(i32), c.POINTER(i32),
                                 np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")],
            "ct_engine_set_row_form": [vp, i32],
            "ct_engine_set_prime_halves": [vp, i32],
            "ct_set_graphs_default": [i32],
            "ct_fsieve_create": [i32, i32, u64, u64, i32, c.POINTER(vp)],
            "ct_fsieve_blocks": [vp, c.POINTER(u64), c.POINTER(u64)],
            "ct_fsieve_set_bmax": [vp, u64],
            "ct_fsieve_run": [vp, u64, u64, u64p, u64, c.POINTER(u64)],
            "ct_fsieve_destroy": [vp],
            "ct_fsieve_check": [i32, i32, u64, u64, i32, c.c_char_p, c.c_double, u64, c.POINTER(FsieveStats)],
            "ct_merge_factors": [i32, i32, u64, u64, c.c_char_p, i32, c.POINTER(c.c_char_p), c.POINTER(i32),
                                 c.POINTER(u64), c.POINTER(u64), c.c_char_p, c.POINTER(MergeStats)],
            "ct_sieve_last_timing": [i32, c.POINTER(SieveTiming)],
            "ct_engine_set_column_split": [vp, i32],
            "ct_engine_set_fused": [vp, i32],
            "ct_prp_begin": [vp, u32p, c.POINTER(vp), c.POINTER(i32), c.POINTER(i32)],
            "ct_prp_finish": [vp, vp, ip, u64p, ip, c.c_void_p],
            "ct_prp_job_free": [vp],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = c.c_int
        L.ct_crc32.argtypes = [u32, c.c_char_p, c.c_size_t]
        L.ct_crc32.restype = u32
        L.ct_cyclo_quit.argtypes = []
        L.ct_cyclo_quit.restype = None
        L.ct_fsieve_quit.argtypes = []
        L.ct_fsieve_quit.restype = None
        L.ct_last_error.argtypes = []
        L.ct_last_error.restype = c.c_char_p
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise CtError(rc, lib().ct_last_error().decode())
    return rc


class Transform:
    """transform + engine of cyclo/src: `lanes` candidates of n = 2^log2_n digits,
    modulo N+ = b^n - b^(n/2) + 1 (sign +1) or N- = b^n - b^(n/2) - 1 (sign -1)."""

    def __init__(self, log2_n, lanes=1, kind=KIND_AUTO, device=0, csize=16, sign=1):
        self.n = 1 << log2_n
        self.log2_n = log2_n
        self.lanes = lanes
        self.sign = sign
        h = ctypes.c_void_p()
        _check(lib().ct_engine_create_ex(device, log2_n, lanes, csize, kind, sign, ctypes.byref(h)))
        self.h = h
        n_ = ctypes.c_uint32(); l_ = ctypes.c_uint32(); k_ = ctypes.c_int()
        _check(lib().ct_engine_info(self.h, ctypes.byref(n_), ctypes.byref(l_), ctypes.byref(k_)))
        self.kind = k_.value
        self.b = None

    def close(self):
        if getattr(self, "h", None):
            lib().ct_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_b(self, b):
        b = np.ascontiguousarray(b, np.uint32)
        assert b.shape == (self.lanes,)
        _check(lib().ct_engine_set_b(self.h, b))
        self.b = b.copy()

    def set(self, a):
        _check(lib().ct_engine_set(self.h, int(a)))

    def copy(self, dst, src):
        _check(lib().ct_engine_copy(self.h, dst, src))

    def square_dup(self, dup=None):
        """dup: per-lane 0/1 sequence (or None)."""
        if dup is None:
            _check(lib().ct_engine_square_dup(self.h, None))
            return
        words = np.zeros((self.lanes + 63) // 64, np.uint64)
        for l, bit in enumerate(dup):
            if bit:
                words[l // 64] |= np.uint64(1) << np.uint64(l % 64)
        _check(lib().ct_engine_square_dup(self.h, words.ctypes.data_as(ctypes.c_void_p)))

    def mul(self, dst=0, a=0, b=1):
        _check(lib().ct_engine_mul_regs(self.h, dst, a, b))

    def get_int(self, reg=0):
        out = np.zeros(self.lanes * self.n, np.uint32)
        _check(lib().ct_engine_get_int(self.h, reg, out))
        return out.reshape(self.lanes, self.n)

    def is_prime(self, canon):
        canon = np.ascontiguousarray(canon, np.uint32).reshape(-1)
        prm = np.zeros(self.lanes, np.int32); r64 = np.zeros(self.lanes, np.uint64)
        rc = lib().ct_engine_is_prime(self.h, canon, prm, r64)
        return rc != 0, prm.astype(bool), r64

    def read_state(self, reg=0):
        out = np.zeros(self.lanes * self.n, np.int32)
        _check(lib().ct_engine_read_state(self.h, reg, out))
        return out.reshape(self.lanes, self.n)

    def write_state(self, reg, digits):
        d = np.ascontiguousarray(digits, np.int32).reshape(-1)
        _check(lib().ct_engine_write_state(self.h, reg, d))

    def set_exponents(self, values):
        """values: one Python int per lane (bit i of lane l drives squaring i)."""
        top = max(int(v).bit_length() for v in values)
        words = max(1, (top + 31) // 32)
        bits = np.zeros((self.lanes, words), np.uint32)
        for l, v in enumerate(values):
            raw = int(v).to_bytes(4 * words, "little")
            bits[l] = np.frombuffer(raw, np.uint32)
        _check(lib().ct_engine_set_exponents(self.h, np.ascontiguousarray(bits.reshape(-1)), words))

    def run(self, i_hi, i_lo=0, gl_block=0):
        _check(lib().ct_engine_run(self.h, int(i_hi), int(i_lo), int(gl_block)))

    def set_column_tile(self, w):
        """Multipass engine: columns per workgroup of the column passes (16 or 32 for n >= 2^13)."""
        _check(lib().ct_engine_set_column_tile(self.h, int(w)))

    def tuning(self):
        """(column tile, split of column pass A, fused column pass, ms per
        squaring for {resident, multipass W=16, W=32, W=16 split, W=32 split,
        W=16 fused, W=32 fused, W=16 fused split, W=32 fused split, the
        fastest with row form 1, with row form 2, resident with two primes one
        after the other, and in workgroup halves}; -1 = not timed, row form
        0/1/2).  For a resident engine the split flag reports the primes in
        workgroup halves.  The row form is last so that the first four
        entries keep their meaning."""
        w, sp, fu, rf = ctypes.c_uint32(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ms = np.zeros(TUNE_TIMINGS, np.float32)
        _check(lib().ct_engine_tuning(self.h, ctypes.byref(w), ctypes.byref(sp), ctypes.byref(fu), ctypes.byref(rf), ms))
        return int(w.value), bool(sp.value), bool(fu.value), [float(x) for x in ms], int(rf.value)

    def set_row_form(self, form):
        """Multipass engine: row pass as one workgroup per (row, prime) (0),
        one workgroup per row transforming its primes in sequence (1), or one
        workgroup per row with a slice of threads per prime (2; used where
        the slices are whole wavefronts within 1024 threads, else 0)."""
        _check(lib().ct_engine_set_row_form(self.h, int(form)))

    def set_prime_halves(self, on):
        """Resident engine: the two primes of a two-prime squaring in the
        halves of one workgroup (transform lengths 2^10 .. 2^13)."""
        _check(lib().ct_engine_set_prime_halves(self.h, int(bool(on))))

    def set_fused(self, on):
        """Multipass engine (N+, n >= 2^13): column pass C fused with the next
        pass A (1, or 2: workgroups always ordered by ticket) or three
        launches per squaring (0)."""
        _check(lib().ct_engine_set_fused(self.h, int(on)))

    def set_column_split(self, on):
        """Multipass engine: column pass A split by prime (1) or combined (0)."""
        _check(lib().ct_engine_set_column_split(self.h, int(bool(on))))

    def context_save(self, path, k, i, elapsed=0.0):
        """Checkpoint registers 0 and 1 in the reference's cyclo.ctx format."""
        _check(lib().ct_context_save(self.h, os.fsencode(path), int(k), int(i), float(elapsed)))

    def context_load(self, path):
        """Restore registers 0 and 1 from a cyclo.ctx file (or its .old); returns (k, i, elapsed)."""
        k, i, el = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(lib().ct_context_load(self.h, os.fsencode(path), ctypes.byref(k), ctypes.byref(i), ctypes.byref(el)))
        return k.value, i.value, el.value

    def set_base(self, base):
        """PRP base (2 or 3): a set exponent bit multiplies by it."""
        _check(lib().ct_engine_set_base(self.h, int(base)))

    def set_graphs(self, on):
        """Multipass engine: replay captured HIP graphs for runs (default) or launch each kernel."""
        _check(lib().ct_engine_set_graphs(self.h, int(bool(on))))

    def set_primes(self, np_):
        """0: automatic (2 primes when the coefficient bound allows), 3: always 3."""
        _check(lib().ct_engine_set_primes(self.h, int(np_)))

    @property
    def primes(self):
        v = ctypes.c_int()
        _check(lib().ct_engine_primes(self.h, ctypes.byref(v)))
        return v.value

    def timer_start(self):
        _check(lib().ct_engine_timer_start(self.h))

    def timer_stop(self):
        ms = ctypes.c_float()
        _check(lib().ct_engine_timer_stop(self.h, ctypes.byref(ms)))
        return ms.value


def prp_check(t, bs, want_residue=False):
    """cyclo::check for one batch: returns (prm, res64, gl_ok, residue or None, status)."""
    bs = np.ascontiguousarray(bs, np.uint32)
    prm = np.zeros(t.lanes, np.int32); r64 = np.zeros(t.lanes, np.uint64); gl = np.zeros(t.lanes, np.int32)
    res = np.zeros(t.lanes * t.n, np.uint32) if want_residue else None
    rc = lib().ct_prp_check(t.h, bs, prm, r64, gl, res.ctypes.data_as(ctypes.c_void_p) if res is not None else None)
    if rc not in (0, -6, -7):
        _check(rc)
    t.b = bs.copy()
    return prm.astype(bool), r64, gl.astype(bool), (res.reshape(t.lanes, t.n) if res is not None else None), rc


class PrpJob:
    """ct_prp_check split at its loop: begin() sets b and the exponent bits
    and gives (i0, gl_block); the caller runs x = 1, d(t) = 1 and the loop
    (Transform.run(i0, 0, gl_block), in pieces); finish() returns what
    prp_check returns."""

    def __init__(self, t, bs):
        self.t = t
        self.bs = np.ascontiguousarray(bs, np.uint32)
        h = ctypes.c_void_p(); i0 = ctypes.c_int(); B = ctypes.c_int()
        _check(lib().ct_prp_begin(t.h, self.bs, ctypes.byref(h), ctypes.byref(i0), ctypes.byref(B)))
        self.h, self.i0, self.gl_block = h, i0.value, B.value
        t.b = self.bs.copy()

    def finish(self, want_residue=False):
        t = self.t
        prm = np.zeros(t.lanes, np.int32); r64 = np.zeros(t.lanes, np.uint64); gl = np.zeros(t.lanes, np.int32)
        res = np.zeros(t.lanes * t.n, np.uint32) if want_residue else None
        rc = lib().ct_prp_finish(t.h, self.h, prm, r64, gl, res.ctypes.data_as(ctypes.c_void_p) if res is not None else None)
        if rc not in (0, -6, -7):
            _check(rc)
        return prm.astype(bool), r64, gl.astype(bool), (res.reshape(t.lanes, t.n) if res is not None else None), rc

    def close(self):
        if getattr(self, "h", None):
            lib().ct_prp_job_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


def residue_checks(canon, bs, qs):
    """Checksums of values x_l = sum_j canon[l, j] b_l^j given by canonical
    digits: RES64 (x mod 2^64, cyclo.h:51-57) and x mod q for each q < 2^31
    (x mod q sees every digit; RES64 sees only the low ones when b is even).
    Returns (res64[lanes], mods[lanes][len(qs)]) as Python ints."""
    d = np.ascontiguousarray(canon, np.uint64).reshape(len(bs), -1)
    lanes, n = d.shape
    b = np.asarray(bs, np.uint64).reshape(-1, 1)
    pw = np.empty((lanes, n), np.uint64); pw[:, 0] = 1
    if n > 1:
        pw[:, 1:] = b
        np.cumprod(pw, axis=1, out=pw)                       # b^j mod 2^64 (wrapping)
    r64 = [int(v) for v in np.sum(d * pw, axis=1, dtype=np.uint64)]
    mods = [[0] * len(qs) for _ in range(lanes)]
    for k, q in enumerate(qs):
        qq = np.uint64(q)
        pq = np.ones((lanes, n), np.uint64)
        cur = b[:, 0] % qq
        step = 1
        while step < n:                                       # pq[:, j] = b^j mod q by doubling
            w = min(step, n - step)
            pq[:, step:step + w] = (pq[:, :w] * cur[:, None]) % qq
            cur = (cur * cur) % qq
            step *= 2
        prod = (d * pq) % qq                                  # digits < 2^31, so products < 2^62
        for l in range(lanes):
            s = 0
            for j in range(0, n, 1 << 15):
                s += int(np.sum(prod[l, j:j + (1 << 15)], dtype=np.uint64))
            mods[l][k] = s % q
    return r64, mods


class Server:
    """Continuous-batching PRP server: submit(id, log2_n, b, steps,
    sign=+1 for N+ or -1 for N-), step() rounds, poll() finished (id, res64,
    hash, latency)."""

    def __init__(self, hbm_budget=8 << 30, slots=64, chunk=250, max_steps=4096, device=0):
        h = ctypes.c_void_p()
        _check(lib().ct_server_create(device, int(hbm_budget), slots, chunk, max_steps, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().ct_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_graphs(self, on):
        _check(lib().ct_server_set_graphs(self.h, int(bool(on))))

    def submit(self, rid, log2_n, b, steps, sign=1):
        if sign == 1:
            _check(lib().ct_server_submit(self.h, int(rid), int(log2_n), int(b), int(steps)))
        else:
            _check(lib().ct_server_submit_ex(self.h, int(rid), int(log2_n), int(b), int(steps), int(sign)))

    def step(self):
        _check(lib().ct_server_step(self.h))

    def poll(self, max_results=4096):
        ids = np.zeros(max_results, np.uint64); r64 = np.zeros(max_results, np.uint64)
        hs = np.zeros(max_results, np.uint64); lat = np.zeros(max_results, np.float64)
        cnt = ctypes.c_uint32()
        _check(lib().ct_server_poll(self.h, ids, r64, hs, lat, max_results, ctypes.byref(cnt)))
        k = cnt.value
        return [(int(ids[i]), int(r64[i]), int(hs[i]), float(lat[i])) for i in range(k)]

    def stats(self):
        q = ctypes.c_uint32(); a = ctypes.c_uint32(); u = ctypes.c_uint64(); s = ctypes.c_uint64()
        _check(lib().ct_server_stats(self.h, ctypes.byref(q), ctypes.byref(a), ctypes.byref(u), ctypes.byref(s)))
        return {"queued": q.value, "active": a.value, "hbm_used": u.value, "squarings": s.value}


def set_graphs_default(on):
    """Graph replay for engines created from now on, the autotuner's included."""
    _check(lib().ct_set_graphs_default(int(bool(on))))


def moduli():
    """The engine's three NTT primes, largest first (host call, no GPU)."""
    p = np.zeros(3, np.uint32)
    _check(lib().ct_engine_moduli(p))
    return [int(x) for x in p]


def res64_string(r):
    """cyclo.h:51-57: uppercase 16-digit hex."""
    return "%016X" % int(r)


def sieve_window(m, b_min, b_max, p_max_pos, p_max_neg, mode=0, p_min_pos=None, p_min_neg=11, device=0,
                 struck=None):
    """Sieve::check(ini) over [b_min, b_max] with explicit p bounds; returns the struck bytes."""
    if p_min_pos is None:
        p_min_pos = 3 * (1 << m) + 1
    if struck is None:
        struck = np.zeros(b_max - b_min + 1, np.uint8)
    _check(lib().ct_sieve_window(device, m, b_min, b_max, p_min_pos, p_max_pos, p_min_neg, p_max_neg, mode, struck))
    return struck


def sieve_last_timing(device=0):
    """Device time of the last sieve_window call on the device: dict with the
    summed duration of the N- root kernel's launches, their number, the
    call's device span (upload to end of download) and its kernels' span
    (between the copies), in ms."""
    t = SieveTiming()
    _check(lib().ct_sieve_last_timing(device, ctypes.byref(t)))
    return {"roots_neg_ms": t.roots_neg_ms, "roots_neg_launches": int(t.roots_neg_launches), "device_ms": t.device_ms,
            "kernels_ms": t.kernels_ms}


def sieve_factors(m, p_min, p_max, mode, device=0, cap=1 << 20):
    """GPU factor emission (ctsieveg): sorted array of (p, b) rows."""
    while True:
        pairs = np.zeros(2 * max(cap, 1), np.uint64)
        cnt = ctypes.c_uint64()
        rc = lib().ct_sieve_factors(device, m, int(p_min), int(p_max), mode, pairs, cap, ctypes.byref(cnt))
        if rc == -3 and cnt.value > cap:
            cap = cnt.value
            continue
        _check(rc)
        out = pairs[:2 * cnt.value].reshape(-1, 2)
        return out[np.lexsort((out[:, 1], out[:, 0]))]


class FactorSieve:
    """Block cursor of the GPU factor emission (gfsieve's per-block driving):
    blocks() = (i_min, i_max); run(i, nblocks) = pairs (p, b) of those blocks,
    sorted."""

    def __init__(self, m, p_min, p_max, mode, device=0):
        h = ctypes.c_void_p()
        _check(lib().ct_fsieve_create(device, m, int(p_min), int(p_max), mode, ctypes.byref(h)))
        self.h = h

    def blocks(self):
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().ct_fsieve_blocks(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_bmax(self, bmax):
        """0: the reference's bMax(p); otherwise pairs with b <= bmax (2^64 - 1: every root)."""
        _check(lib().ct_fsieve_set_bmax(self.h, int(bmax)))

    def run(self, i_begin, nblocks, cap=1 << 20):
        while True:
            pairs = np.zeros(2 * max(cap, 1), np.uint64)
            cnt = ctypes.c_uint64()
            rc = lib().ct_fsieve_run(self.h, int(i_begin), int(nblocks), pairs, cap, ctypes.byref(cnt))
            if rc == -3 and cnt.value > cap:
                cap = cnt.value
                continue
            _check(rc)
            out = pairs[:2 * cnt.value].reshape(-1, 2)
            return out[np.lexsort((out[:, 1], out[:, 0]))]

    def close(self):
        if getattr(self, "h", None):
            lib().ct_fsieve_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def fsieve_check(m, p_min, p_max, mode, work_dir, record_s=120.0, max_blocks=0, device=0):
    """gfsieve::check on the device: fp/fn<ext> factor files and the ctx<ext>
    checkpoint in work_dir; returns (status, stats dict), status 0 = done,
    INTERRUPTED = stopped with the checkpoint written."""
    st = FsieveStats()
    rc = lib().ct_fsieve_check(device, m, int(p_min), int(p_max), mode, os.fsencode(work_dir), float(record_s),
                               int(max_blocks), ctypes.byref(st))
    if rc not in (0, INTERRUPTED):
        _check(rc)
    return rc, {f: getattr(st, f) for f, _ in FsieveStats._fields_}


def fsieve_quit():
    lib().ct_fsieve_quit()


def merge_factors(m, b_min, b_max, cand_path, factor_files, out_path, device=0):
    """The ctools merge: remove from the candidate list every b with a
    verified factor.  factor_files: [(path, mode, p_min_t, p_max_t)] with
    mode +1 (fp file) or -1 (fn file) and the file's p range in units of
    10^12 (None: no range check).  Returns the stats dict."""
    k = len(factor_files)
    paths = (ctypes.c_char_p * max(k, 1))(*[os.fsencode(f[0]) for f in factor_files])
    modes = (ctypes.c_int * max(k, 1))(*[int(f[1]) for f in factor_files])
    ranged = k > 0 and all(f[2] is not None and f[3] is not None for f in factor_files)
    lo = (ctypes.c_uint64 * max(k, 1))(*[int(f[2]) if ranged else 0 for f in factor_files])
    hi = (ctypes.c_uint64 * max(k, 1))(*[int(f[3]) if ranged else 0 for f in factor_files])
    st = MergeStats()
    _check(lib().ct_merge_factors(device, m, int(b_min), int(b_max), os.fsencode(cand_path), k, paths, modes,
                                  lo if ranged else None, hi if ranged else None, os.fsencode(out_path),
                                  ctypes.byref(st)))
    return {f: getattr(st, f) for f, _ in MergeStats._fields_}


def crc32(data, crc=0):
    """CRC-32 (IEEE) of bytes, continued from crc."""
    return int(lib().ct_crc32(crc, bytes(data), len(data)))


def sieve_save(directory, m, b_min, b_max, p_min_pos, p_min_neg, struck, write_cand=True):
    """Sieve::write: sv_<m>_<b_min>_<b_max>.dat (and cand_*.txt) in directory."""
    struck = np.ascontiguousarray(struck, np.uint8)
    _check(lib().ct_sieve_save(os.fsencode(directory), m, b_min, b_max, p_min_pos, p_min_neg, struck, int(write_cand)))


def sieve_load(directory, m, b_min, b_max):
    """Sieve::read: (struck bytes, p_min_pos, p_min_neg) from sv_<m>_<b_min>_<b_max>.dat."""
    struck = np.zeros(b_max - b_min + 1, np.uint8)
    pp, pn = ctypes.c_uint64(), ctypes.c_uint64()
    _check(lib().ct_sieve_load(os.fsencode(directory), m, b_min, b_max, ctypes.byref(pp), ctypes.byref(pn), struck))
    return struck, pp.value, pn.value


def cyclo_check_file(log2_n, cand_file, work_dir, lanes=64, checkpoint_s=600.0, max_steps=0, device=0):
    """cyclo::check(filename) on the device: returns (status, stats dict); status
    0 = done, INTERRUPTED = stopped with cyclo.ctx saved."""
    st = CycloStats()
    rc = lib().ct_cyclo_check_file(device, log2_n, os.fsencode(cand_file), os.fsencode(work_dir), lanes,
                                   float(checkpoint_s), int(max_steps), ctypes.byref(st))
    if rc not in (0, INTERRUPTED):
        _check(rc)
    return rc, {f: getattr(st, f) for f, _ in CycloStats._fields_}


def cyclo_quit():
    lib().ct_cyclo_quit()


def survivors(struck, b_min):
    return np.nonzero(struck == 0)[0].astype(np.uint64) + np.uint64(b_min)
