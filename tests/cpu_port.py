"""ctypes binding of the C++ CPU comparator (oracle/libctcpu.so) -- test
infrastructure and the bench's CPU baseline; the product never uses it."""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "libctcpu.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "libctcpu.so"])
        L = ctypes.CDLL(LIB_PATH)
        u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        L.cp_new.argtypes = [ctypes.c_uint32, ctypes.c_uint32, u32p]
        L.cp_new.restype = ctypes.c_void_p
        L.cp_free.argtypes = [ctypes.c_void_p]
        L.cp_steps.argtypes = [ctypes.c_void_p, i32p, u32p, ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.cp_steps.restype = ctypes.c_int
        _lib = L
    return _lib


class CpuEngine:
    """`lanes` candidates of n = 2^ln digits (two-prime mode only)."""

    def __init__(self, ln, bs):
        self.ln, self.n, self.lanes = ln, 1 << ln, len(bs)
        self.h = lib().cp_new(ln, len(bs), np.ascontiguousarray(bs, np.uint32))
        if not self.h:
            raise ValueError("the CPU comparator needs two-prime sizes (3 n b^2 < p1 p2 / 2)")

    def steps(self, digits, bits, words, i_hi, nsteps, threads=1):
        assert digits.dtype == np.int32 and digits.flags.c_contiguous
        rc = lib().cp_steps(self.h, digits, np.ascontiguousarray(bits, np.uint32), words, i_hi, nsteps, threads)
        assert rc == 0

    def close(self):
        if self.h:
            lib().cp_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


def pack_bits(values, lanes):
    top = max(int(v).bit_length() for v in values)
    words = max(1, (top + 31) // 32)
    bits = np.zeros((lanes, words), np.uint32)
    for l, v in enumerate(values):
        bits[l] = np.frombuffer(int(v).to_bytes(4 * words, "little"), np.uint32)
    return bits.reshape(-1), words
