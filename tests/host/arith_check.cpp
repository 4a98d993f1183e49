// Host check of the device arithmetic in ctwin_amd/csrc/ct_arith.h (the same
// header, compiled by g++): exact truncating division by b against __int128,
// the biased division for |a| < 2^62, the lazy Shoup and Montgomery products
// against their definitions.
#include <cstdio>
#include <cstdlib>
#include <random>

#include "../../ctwin_amd/csrc/ct_arith.h"

using namespace ct;

int main(int argc, char **argv)
{
	const long iters = (argc > 1) ? std::atol(argv[1]) : 2000000;
	std::mt19937_64 rng(12345);
	long bad = 0;
	const uint32_t bs[] = {33, 34, 65537, 73764, 327968, 1048575, (1u << 30) + 9699690u, 0x7fffffffu};
	for (long it = 0; it < iters; ++it)
	{
		const uint32_t b = (it % 3 == 0) ? bs[it % 8] : (uint32_t)(33 + rng() % (0x7fffffffu - 33));
		const divisor dv = make_divisor(b);
		int64_t a;
		switch (it % 4)
		{
		case 0: a = (int64_t)(rng() >> 1); break;                                   // up to 2^63 - 1
		case 1: a = -(int64_t)(rng() >> 1); break;
		case 2: a = (int64_t)(rng() % (1ull << 40)) - (1ll << 39); break;
		default: a = (int64_t)((rng() % 5) * (uint64_t)b) - 2 * (int64_t)b + (int64_t)(rng() % 3) - 1; break;
		}
		int64_t x = a;
		const int32_t r = divrem_i64(x, dv);
		const __int128 q = (__int128)a / b, rr = (__int128)a % b;      // C truncation
		if (x != (int64_t)q || r != (int32_t)rr) { if (bad++ < 5) std::printf("div a=%lld b=%u\\n", (long long)a, b); }
		// the biased form, for |a| < 2^62 - b: random values, the edges of the
		// range, values near multiples of b
		int64_t a62;
		const int64_t edge = ((int64_t)1 << 62) - (int64_t)b - 1;
		switch (it % 5)
		{
		case 0: a62 = a % ((int64_t)1 << 61); break;
		case 1: a62 = edge - (int64_t)(rng() % 1000); break;
		case 2: a62 = -(edge - (int64_t)(rng() % 1000)); break;
		case 3: a62 = (int64_t)(rng() % 1000 + 1) * (int64_t)b * ((rng() & 1) ? 1 : -1) + (int64_t)(rng() % 3) - 1; break;
		default: a62 = a % ((int64_t)1 << 40); break;
		}
		int64_t y = a62;
		const int32_t r62 = divrem_i62(y, dv);
		const __int128 q62 = (__int128)a62 / b, rr62 = (__int128)a62 % b;
		if (y != (int64_t)q62 || r62 != (int32_t)rr62) { if (bad++ < 5) std::printf("div62 a=%lld b=%u\n", (long long)a62, b); }
	}
	for (int P = 0; P < kPrimes; ++P)
	{
		const uint32_t p = kP[P];
		for (long it = 0; it < iters / 4; ++it)
		{
			const uint32_t w = (uint32_t)(rng() % p), a = (uint32_t)rng();
			const uint32_t wp = (uint32_t)(((unsigned __int128)w << 32) / p);
			const uint32_t t = shoup(a, w, wp, p);
			if (t >= 2 * p || t % p != (uint32_t)(((uint64_t)a * w) % p)) { if (bad++ < 5) std::printf("shoup\\n"); }
			const uint32_t sn = shoup_neg(a, 0u - w, wp, p);
			if ((uint32_t)(0u - sn) != t) { if (bad++ < 5) std::printf("shoup_neg\\n"); }
			const uint32_t x = (uint32_t)(rng() % (2ull * p)), y = (uint32_t)(rng() % (2ull * p));
			const uint32_t m = mont_lazy(x, y, p, kQ[P]);
			// m * 2^32 = x y (mod p), m in (0, 2p)
			if (m == 0 || m >= 2 * p || (((uint64_t)m << 32) % p) != ((uint64_t)x * y) % p) { if (bad++ < 5) std::printf("mont\\n"); }
		}
	}
	std::printf("%ld mismatches\n", bad);
	return bad ? 1 : 0;
}
