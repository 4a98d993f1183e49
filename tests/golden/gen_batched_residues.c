/*
 * Fixture generator (test infrastructure): GMP reference residues of the
 * base-2 PRP loop of a batch of candidates run in lockstep, as one engine
 * launch runs them: x <- x^2 * 2^(bit i of E) mod N, N = b^n - b^(n/2) + 1,
 * E = N - 1, i from a common top bit index i0 (the largest over the batch;
 * a lane whose E is shorter sees leading zero bits) down, x = 1 initially
 * (cyclo.h:277-308).  For every b, prints one JSON line with RES64 = x mod 2^64
 * and x mod three fixed primes q < 2^31 at each checkpoint t (squarings done).
 * x mod q pins every digit; RES64 alone sees only the low digits when b is
 * even (b^j = 0 mod 2^64 once j v2(b) >= 64).
 *
 * build: gcc -O2 -I/opt/conda/include gen_batched_residues.c -L/opt/conda/lib -lgmp -o gen_batched
 * run:   ./gen_batched [-m] <log2_n> <i0> <t1,t2,...> <b> [<b> ...]
 * -m: the other twin member, N = b^n - b^(n/2) - 1 (E = N - 1 likewise).
 */
#include <gmp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int main(int argc, char **argv)
{
	int minus = 0;
	if (argc > 1 && strcmp(argv[1], "-m") == 0) { minus = 1; ++argv; --argc; }
	if (argc < 5) { fprintf(stderr, "usage: gen_batched [-m] log2_n i0 t1,t2,... b [b ...]\n"); return 1; }
	const unsigned ln = (unsigned)atoi(argv[1]);
	const long i0 = atol(argv[2]);
	long ts[256];
	int nt = 0;
	for (char *s = strtok(argv[3], ","); s && nt < 256; s = strtok(NULL, ",")) ts[nt++] = atol(s);
	const unsigned long n = 1ul << ln;
	const unsigned long q[3] = {2147483629ul, 2147483587ul, 2147483579ul};
	mpz_t h, N, E, x;
	mpz_inits(h, N, E, x, NULL);
	for (int a = 4; a < argc; ++a)
	{
		const unsigned long b = strtoul(argv[a], 0, 10);
		mpz_ui_pow_ui(h, b, n / 2);
		mpz_mul(E, h, h);
		mpz_sub(E, E, h);
		if (minus) mpz_sub_ui(E, E, 2);
		mpz_add_ui(N, E, 1);
		if ((long)mpz_sizeinbase(E, 2) - 1 > i0) { fprintf(stderr, "b=%lu: E longer than i0\n", b); return 2; }
		mpz_set_ui(x, 1);
		printf("{\"b\": %lu, \"res64\": [", b);
		int k = 0;
		for (long t = 1; k < nt; ++t)
		{
			const long i = i0 - (t - 1);
			mpz_mul(x, x, x);
			if (i >= 0 && mpz_tstbit(E, (mp_bitcnt_t)i)) mpz_mul_2exp(x, x, 1);
			mpz_mod(x, x, N);
			if (t == ts[k])
			{
				printf("%s\"%016lX\"", k ? ", " : "", (unsigned long)mpz_getlimbn(x, 0));
				static unsigned long mq[256][3];
				for (int j = 0; j < 3; ++j) mq[k][j] = mpz_fdiv_ui(x, q[j]);
				++k;
				if (k == nt)
				{
					printf("], \"mod_q\": [");
					for (int c = 0; c < nt; ++c)
						printf("%s[%lu, %lu, %lu]", c ? ", " : "", mq[c][0], mq[c][1], mq[c][2]);
					printf("]}\n");
				}
			}
		}
		fflush(stdout);
	}
	mpz_clears(h, N, E, x, NULL);
	return 0;
}
