// Drives the transform binding (cyclo_binding.h) the way cyclo::check drives
// the reference's transform (cyclo/src/cyclo.h:252-308): register 0 := 1,
// register 1 := 1, then for every exponent bit from the top a squareDup with
// the lanes' bits, and at the Gerbicz-Li block boundaries d(t) := d(t) x
// (copy 2 <- 0, mul, copy 1 <- 0, copy 0 <- 2); finally getInt + isPrime.
// It prints one line per lane: lane, b, prm, RES64 of x = 2^(N-1) mod N, and
// RES64 of x^2 mod N (one more copy + mul), for a checker to compare with GMP.
//
//   cyclo_binding <log2_n> <b_0> ... <b_k>      (k < 64; later lanes repeat b_k,
//                                                 as cyclo.h:216-234 fills a batch)
//   cyclo_binding --check                       (link check: no device call)
#include <gmp.h>

#include <cinttypes>
#include <cstring>

#include "cyclo_binding_prelude.h"
#include "cyclo_binding.h"

static uint64 lane_bits(const std::vector<mpz_t *> &e, size_t i)
{
	uint64 d = 0;
	for (size_t j = 0; j < VSIZE; ++j) d |= uint64(mpz_tstbit(*e[j], mp_bitcnt_t(i)) != 0) << j;
	return d;
}

int main(int argc, char **argv)
{
	if (argc == 2 && !strcmp(argv[1], "--check"))
	{
		const char *msg = ct_last_error();                  // the library loads and resolves; no device call
		printf("ok%s%s\n", (msg && *msg) ? ": " : "", msg ? msg : "");
		return 0;
	}
	if (argc < 3) { fprintf(stderr, "usage: %s <log2_n> <b_0> ... | --check\n", argv[0]); return 2; }
	const int ln = atoi(argv[1]);
	const size_t n = size_t(1) << ln;
	vint32 b;
	for (size_t i = 0; i < VSIZE; ++i) b[i] = uint32(strtoul(argv[2 + std::min<size_t>(i, size_t(argc - 3))], nullptr, 10));

	// exponents N - 1 = b^n - b^(n/2) per lane
	std::vector<mpz_t *> e(VSIZE);
	int i0 = 0;
	for (size_t j = 0; j < VSIZE; ++j)
	{
		e[j] = new mpz_t[1];
		mpz_t t; mpz_init(t);
		mpz_init(*e[j]);
		mpz_ui_pow_ui(*e[j], b[j], n);
		mpz_ui_pow_ui(t, b[j], n / 2);
		mpz_sub(*e[j], *e[j], t);
		mpz_clear(t);
		i0 = std::max(i0, int(mpz_sizeinbase(*e[j], 2)) - 1);
	}

	engine eng;
	transform tr(n, eng, false, 16);
	tr.set_b(b);
	tr.set(1);
	tr.copy(1, 0);
	const int L = 2 << (ilog2_32(uint32_t(i0)) / 2), B_GL = int((i0 - 1) / L) + 1;
	for (int i = i0; i >= 0; --i)
	{
		tr.squareDup(lane_bits(e, size_t(i)));
		if ((i % B_GL == 0) && (i / B_GL != 0))
		{
			tr.copy(2, 0);
			tr.mul();
			tr.copy(1, 0);
			tr.copy(0, 2);
		}
	}
	tr.getInt(0);
	bool prm[VSIZE];
	uint64_t r64[VSIZE], r64sq[VSIZE];
	if (tr.isPrime(prm, r64)) { fprintf(stderr, "computation failed\n"); return 1; }
	tr.copy(1, 0);
	tr.mul();
	tr.getInt(0);
	bool prm2[VSIZE];
	if (tr.isPrime(prm2, r64sq)) { fprintf(stderr, "computation failed\n"); return 1; }
	for (size_t j = 0; j < VSIZE; ++j)
		printf("%zu %u %d %016" PRIX64 " %016" PRIX64 "\n", j, b[j], prm[j] ? 1 : 0, r64[j], r64sq[j]);
	for (size_t j = 0; j < VSIZE; ++j) { mpz_clear(*e[j]); delete[] e[j]; }
	return 0;
}
