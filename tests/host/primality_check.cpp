// Host check of the 32-bit strong probable-prime test spsp_2_7_61 of
// ctwin_amd/csrc/ct_arith.h (the same header, compiled by g++): exact against
// a sieve of Eratosthenes for every odd p in [43, 2^22), against trial division
// for random 32-bit values and for values at the edges of the ranges
// (2^31, 2^32), and on strong pseudoprimes to some of the bases.
#include <cstdio>
#include <random>
#include <vector>

#include "../../ctwin_amd/csrc/ct_arith.h"

using namespace ct;

static bool trial(uint32_t n, const std::vector<uint32_t> &small)
{
	if (n < 2) return false;
	for (uint32_t p : small)
	{
		if ((uint64_t)p * p > n) return true;
		if (n % p == 0) return n == p;
	}
	return true;
}

int main()
{
	long bad = 0, checked = 0;
	const uint32_t L = 1u << 22;
	std::vector<uint8_t> comp(L, 0);
	for (uint32_t i = 2; (uint64_t)i * i < L; ++i)
		if (!comp[i])
			for (uint32_t j = i * i; j < L; j += i) comp[j] = 1;
	std::vector<uint32_t> small;
	for (uint32_t i = 2; i < 65536; ++i) if (!comp[i]) small.push_back(i);
	for (uint32_t p = 43; p < L; p += 2)
	{
		++checked;
		if (spsp_2_7_61(p) != !comp[p]) { if (bad++ < 5) std::printf("sieve p=%u\n", p); }
	}
	std::vector<uint32_t> vals = {2047u, 3277u, 4033u, 4681u, 8321u, 314821u, 3215031751u, 25326001u, 161304001u,
	                              960946321u, 1157839381u, 2147483647u, 2147483649u, 2147483659u, 4294967291u,
	                              4294967295u, 4294967279u, 4294836225u /* 65535^2 */, 4294967297u & 0xffffffffu};
	for (uint32_t d = 1; d < 2000; d += 2) { vals.push_back(0x80000000u - d); vals.push_back(0x80000000u + d); vals.push_back(0u - d); }
	std::mt19937 rng(2024);
	for (int i = 0; i < 300000; ++i) vals.push_back(rng() | 1u);
	for (uint32_t v : vals)
	{
		if (v < 43 || (v & 1) == 0) continue;
		++checked;
		if (spsp_2_7_61(v) != trial(v, small)) { if (bad++ < 5) std::printf("value %u\n", v); }
	}
	std::printf("%ld mismatches in %ld checks\n", bad, checked);
	return bad ? 1 : 0;
}
