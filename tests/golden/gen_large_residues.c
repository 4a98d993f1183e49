/*
 * Fixture generator (test infrastructure): GMP reference residues of the
 * base-2 PRP loop x <- x^2 * 2^(bit i of E) mod N, N = b^n - b^(n/2) + 1,
 * E = N - 1, i from the top bit down, x = 1 initially (cyclo.h:277-308).
 * Prints one JSON line per checkpoint: step t, RES64 = x mod 2^64 and x mod q
 * for three fixed primes q < 2^31.
 *
 * build: gcc -O2 -I/opt/conda/include gen_large_residues.c -L/opt/conda/lib -lgmp -o gen
 * run:   ./gen <log2_n> <b> <steps> <every>
 */
#include <gmp.h>
#include <stdio.h>
#include <stdlib.h>

int main(int argc, char **argv)
{
	if (argc != 5) { fprintf(stderr, "usage: gen log2_n b steps every\n"); return 1; }
	const unsigned ln = (unsigned)atoi(argv[1]);
	const unsigned long b = strtoul(argv[2], 0, 10);
	const long steps = atol(argv[3]), every = atol(argv[4]);
	const unsigned long n = 1ul << ln;
	const unsigned long q[3] = {2147483629ul, 2147483587ul, 2147483579ul};
	mpz_t h, N, E, x;
	mpz_inits(h, N, E, x, NULL);
	mpz_ui_pow_ui(h, b, n / 2);
	mpz_mul(E, h, h);
	mpz_sub(E, E, h);
	mpz_add_ui(N, E, 1);
	const long i0 = (long)mpz_sizeinbase(E, 2) - 1;
	mpz_set_ui(x, 1);
	printf("{\"log2_n\": %u, \"b\": %lu, \"i0\": %ld, \"q\": [%lu, %lu, %lu]}\n", ln, b, i0, q[0], q[1], q[2]);
	fflush(stdout);
	for (long t = 1; t <= steps; ++t)
	{
		const long i = i0 - (t - 1);
		mpz_mul(x, x, x);
		if (mpz_tstbit(E, (mp_bitcnt_t)i)) mpz_mul_2exp(x, x, 1);
		mpz_mod(x, x, N);
		if (t % every == 0)
		{
			const unsigned long lo = mpz_getlimbn(x, 0);
			printf("{\"t\": %ld, \"res64\": \"%016lX\", \"mod_q\": [%lu, %lu, %lu]}\n", t, lo,
			       mpz_fdiv_ui(x, q[0]), mpz_fdiv_ui(x, q[1]), mpz_fdiv_ui(x, q[2]));
			fflush(stdout);
		}
	}
	mpz_clears(h, N, E, x, NULL);
	return 0;
}
