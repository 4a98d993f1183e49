"""The device arithmetic header compiled for the host (g++) and checked against
exact 128-bit integer arithmetic: truncating division by b (the carry), lazy
Shoup products, lazy Montgomery products."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_device_arithmetic_on_host(tmp_path):
    exe = str(tmp_path / "arith_check")
    src = os.path.join(HERE, "host", "arith_check.cpp")
    r = subprocess.run(["g++", "-O2", "-std=c++17", src, "-o", exe], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stderr)
    out = subprocess.run([exe, "1000000"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout
    assert out.stdout.strip().startswith("0 mismatches")


def test_32bit_primality_on_host(tmp_path):
    """The sieve's 32-bit primality test (strong probable primes to 2, 7, 61 in
    Montgomery form) against exact primality: every odd p < 2^22, random
    32-bit values, the edges of the ranges, strong pseudoprimes."""
    exe = str(tmp_path / "primality_check")
    src = os.path.join(HERE, "host", "primality_check.cpp")
    r = subprocess.run(["g++", "-O2", "-std=c++17", src, "-o", exe], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stderr)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout
    assert out.stdout.strip().startswith("0 mismatches")


def test_sieve_roots_on_host(tmp_path):
    """The sieve's 2-Sylow generator (least non-residue by reciprocity) and
    2^d-th roots against the Euler-criterion scan and the direct construction
    with Fermat inversions, and against x^(2^d) = y: every prime 10k +- 1
    below 2^18, primes 1 mod 2^12 below 2^26, the Fermat primes, random
    32-bit and 62-bit primes."""
    exe = str(tmp_path / "numth_check")
    src = os.path.join(HERE, "host", "numth_check.cpp")
    r = subprocess.run(["g++", "-O2", "-std=c++17", src, "-o", exe], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stderr)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout
    assert out.stdout.strip().startswith("0 mismatches")
