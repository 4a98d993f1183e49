/*
 * ctwin oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the two data-parallel paths of galloty/ctwin, used
 * as the checker by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py.  The product library (ctwin_amd) never links, loads or calls this
 * code; its GPU path fails loudly when its HIP library is missing.
 *
 * Path A -- exact Fermat squaring engine for N = b^n - b^(n/2) + 1, n = 2^m:
 *   the reference's transform (cyclo/src/transform.h, cyclo/ocl/kernel.cl,
 *   cyclo/algo/algo.cpp) restated with plain modular arithmetic: the root
 *   table, the Phi_6 split, the sqr2-fused innermost level, the inverse merge,
 *   Garner CRT over three primes, the chunked signed base-b carry
 *   (normalize1/normalize2), the canonical read-out, RES64 and the
 *   Gerbicz-Li PRP schedule of cyclo/src/cyclo.h.
 * Path B -- the b-window sieve of sieve/CPU/ctsieve.cpp (+ and - sides).
 *
 * Pinning (each function is checked by tests/test_oracle_*.py against data or
 * mathematics that a plausible mistake would fail):
 *   or_transform_tables   root identities w[2s+2j]^2 = w[s+j],
 *                         w[2s+2j+1]^2 = -w[s+j], w*wi = 1, r^2 - r + 1 = 0
 *   or_poly_sqr_algo      the PARI/GP known-answer strings of
 *                         cyclo/algo/algo.cpp:175-176
 *   or_square_mod/or_mul_mod  exact schoolbook product mod X^n - X^(n/2) + 1
 *                         (or_poly_mul_exact) on random inputs, every prime
 *   or_poly_mul_exact     Python big-integer polynomial product reduced by
 *                         X^n = X^(n/2) - 1 (independent code)
 *   or_garner3            Python integer CRT on random residues
 *   or_engine_square_dup/or_engine_mul  value mod N equals Python
 *                         (x^2 * 2^dup) mod N / (x*y) mod N on random digits
 *   or_get_int/or_is_prime Python int(x) mod N digits, RES64 = value mod 2^64
 *   or_exponent           Python b**n - b**(n/2)
 *   or_prp                Python pow(2, N-1, N) residues, known PRPs
 *                         (result 1), Gerbicz-Li pass and forced failure
 *   or_sieve_window       brute force over every prime p and b (tiny windows)
 *                         and the survivor lists written by the reference
 *                         program compiled from its own source (oracle/_ref)
 */
#ifndef CTWIN_ORACLE_H
#define CTWIN_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- path A: squaring engine ---------------- */

uint32_t or_powmod32(uint32_t a, uint64_t e, uint32_t p);

/* Root tables of transform.h:336-366 in plain (non-Montgomery) residues.
 * w[s + j] for s = 1, 2, ..., n/2; wi[s + j] = -w[s + (s - 1 - j)];
 * wi[0] = w[1]; wi[1] = 2 (1 - 2 r)^-1 * (2 / n). */
void or_transform_tables(uint32_t p, uint32_t g, uint32_t n, uint32_t *w, uint32_t *wi);

/* x <- x^2 mod (X^n - X^(n/2) + 1, p), the squareDup transform of
 * engine.h:387-419 (forward levels, sqr2, inverse levels, merge). */
void or_square_mod(uint32_t p, uint32_t n, const uint32_t *w, const uint32_t *wi, uint32_t *x);
/* x <- x * y mod (X^n - X^(n/2) + 1, p), engine.h:421-433: full forward of
 * both, pointwise product halved, full inverse.  y is preserved. */
void or_mul_mod(uint32_t p, uint32_t n, const uint32_t *w, const uint32_t *wi, uint32_t *x, const uint32_t *y);
/* The single-prime CPU transform of cyclo/algo/algo.cpp:87-141 (P1, g = 19). */
void or_poly_sqr_algo(uint32_t n, uint32_t *x);
/* Exact schoolbook product c = a * b mod X^n - X^(n/2) + 1 over Z. */
void or_poly_mul_exact(uint32_t n, const int64_t *a, const int64_t *b, __int128 *c);

/* Garner CRT of kernel.cl:451-460, generic primes, result in (-M/2, M/2]. */
__int128 or_garner3(uint32_t r1, uint32_t r2, uint32_t r3, uint32_t p1, uint32_t p2, uint32_t p3);
/* Same, split into (lo, hi) 64-bit words for callers without __int128. */
void or_garner3_words(uint32_t r1, uint32_t r2, uint32_t r3, uint32_t p1, uint32_t p2, uint32_t p3,
                      uint64_t *lo, int64_t *hi);

/* normalize1 + normalize2 of kernel.cl:1029-1081 for one candidate:
 * coef[n] exact coefficients, dup doubles them, csize-digit chunks. */
void or_normalize(uint32_t n, uint32_t csize, uint32_t b, int dup, const __int128 *coef, int32_t *digits);

/* A squaring context: three primes p[3] with generators g[3], length n,
 * chunk size csize (2 <= csize <= n/2) and base b (32 < b < 2^31). */
void *or_engine_new(const uint32_t *p, const uint32_t *g, uint32_t n, uint32_t csize, uint32_t b);
void or_engine_free(void *e);
/* One squareDup step on the signed digit state: x <- x^2 * 2^dup. */
void or_engine_square_dup(void *e, int dup, int32_t *digits);
/* x <- x * y (the Gerbicz-Li product), carried without doubling. */
void or_engine_mul(void *e, int32_t *x, const int32_t *y);
/* The exact coefficients (before the carry) of x^2, as (lo, hi) words. */
void or_engine_square_coefs(void *e, const int32_t *digits, uint64_t *lo, int64_t *hi);

/* Canonical digits in [0, b) of the value mod N (transform.h:501-597).
 * Returns 0, or -1 when the reference would report "getInt failed". */
int or_get_int(uint32_t n, uint32_t b, const int32_t *digits, uint32_t *canon);
/* transform.h:600-636: *prm = (value == 1), *res64 = value mod 2^64,
 * returns 1 when the value is 0 (computation failed). */
int or_is_prime(uint32_t n, uint32_t b, const uint32_t *canon, int *prm, uint64_t *res64);

/* Exponent E = b^n - b^(n/2) as little-endian 32-bit words; returns the word
 * count.  The buffer must hold n + 2 words (b < 2^32). */
size_t or_exponent(uint32_t n, uint32_t b, uint32_t *words);

/* Full base-2 Fermat PRP test with Gerbicz-Li (cyclo.h:252-411) for one
 * candidate.  i0 = top bit index used for the schedule (the reference takes
 * the maximum over its lanes); pass 0 to use this candidate's own.
 * flip_at >= 0 flips the exponent bit at that iteration (the commented fault
 * hook of cyclo.h:298) so that Gerbicz-Li must fail.
 * Outputs: canonical residue canon[n], *res64, *prm, *gl_ok.
 * Returns 0, 1 if the residue is 0 ("computation failed"), -1 on a
 * read-out failure. */
int or_prp(void *e, int i0, int flip_at, uint32_t *canon, uint64_t *res64, int *prm, int *gl_ok);

/* The PRP loop alone: x <- 1, then for `steps` iterations
 * x <- x^2 * 2^(bit i of E) with i = i0, i0 - 1, ...  (i0 = 0: own top bit).
 * The signed digit state is returned. */
void or_prp_steps(void *e, int i0, int steps, int32_t *digits);

/* Same loop with the exponent bits given as a window: bit i of E is
 * (win[(i - i_lo_win) / 32] >> ((i - i_lo_win) % 32)) & 1 for i >= i_lo_win
 * (for sizes where computing E in full would be too slow). */
void or_prp_steps_window(void *e, const uint32_t *win, int i_lo_win, int i0, int steps, int32_t *digits);

/* ---------------- path B: window sieve ---------------- */

int or_jacobi(uint64_t x, uint64_t y);
int or_isprime64(uint64_t n);

/* Sieve::check_pos / check_neg / check_root of ctsieve.cpp:653-787.
 * struck[b - b_min] (bytes) is read and updated.  mode: +1 (+ side only),
 * -1 (- side only), 0 (both).  p_min_* follow the reference's _p_min_*
 * (ini: 3*2^m + 1 and 11).  Returns 0, or -1 on "Calculation error". */
int or_sieve_window(int m, uint64_t b_min, uint64_t b_max, uint64_t p_min_pos, uint64_t p_max_pos,
                    uint64_t p_min_neg, uint64_t p_max_neg, int mode, uint8_t *struck);

/* GPU factor emission of sieve/GPU (kernel.cl:243-479 + ctsieve.h:248-297)
 * restated on the CPU: for p in [p_min, p_max] of the side's form that pass
 * the base-2 Euler test, the root seeds of init_factors_* (a-search with the
 * Kronecker symbol, giving up at a >= 256; Atkin / Cipolla square roots), the
 * root enumeration of generate_factors_* with b <= bMax (5e9, or 1e10 above
 * 1e16 (+) / 1e13 (-)), and the host re-verification of saveFactors (pairs
 * that fail are dropped when p is composite; -1 is returned when p is prime).
 * Pairs are written as pairs[2i] = p, pairs[2i+1] = b, in emission order;
 * *count gets the total even when it exceeds cap. */
int or_sieve_factors(int m, uint64_t p_min, uint64_t p_max, int mode, uint64_t *pairs, uint64_t cap, uint64_t *count);

/* Work counters of the two functions above: 64-bit Montgomery products
 * (MpArith::mul, ctsieve.cpp:175-395 / kernel.cl's Montgomery mul) and the
 * products of the strike and pair verifications.  reset != 0 clears them after
 * reading.  Not thread-safe: a measurement aid for single-threaded runs. */
void or_sieve_products(int reset, uint64_t *mont, uint64_t *verify);

#ifdef __cplusplus
}
#endif

#endif
