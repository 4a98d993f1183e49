// Host-side root tables for the transform modulo X^n - X^(n/2) + 1.
//
// The factorisation tree follows cyclo/src/transform.h:336-366: level 1 splits
// X^n - X^(n/2) + 1 = (X^(n/2) - r)(X^(n/2) - r') with r a primitive 6th root of
// unity (r + r' = 1); below it, block j of level s (s blocks) is reduced by its
// root w[s + j] = r_s^(3 bitrev_s(j) + 1 + floor(2j/s)), r_s = g^((p-1)/(6s)).
// The inverse table holds w[s + j]^-1.  Entry 0/1 of the inverse table carry
// the level-1 merge constants.
//
// Table layout (uint2 entries = (w * 2^32 mod p, that value * p^-1 mod 2^32)):
//   tw[((P * 2) + dir) * n + idx], P = prime, dir 0 = forward, 1 = inverse.
#pragma once

#include <cstdint>
#include <vector>

#include "ct_arith.h"

namespace ct {

struct tw_entry { uint32_t w, wq; };

inline uint32_t bitrev_n(uint32_t i, uint32_t n)
{
	uint32_t r = 0;
	for (uint32_t k = n, j = i; k > 1; k /= 2, j /= 2) r = (2 * r) | (j % 2);
	return r;
}

inline tw_entry make_entry(uint32_t w_mont, int P)
{
	tw_entry e; e.w = w_mont; e.wq = w_mont * kQ[P];
	return e;
}

// Returns 3 * 2 * n entries.
inline std::vector<tw_entry> build_tables(uint32_t n)
{
	std::vector<tw_entry> tw(6 * (size_t)n);
	for (int P = 0; P < kPrimes; ++P)
	{
		const uint32_t p = kP[P], g = kG[P];
		tw_entry *fw = &tw[(size_t)(P * 2 + 0) * n];
		tw_entry *iw = &tw[(size_t)(P * 2 + 1) * n];
		uint32_t r = 0;
		for (uint32_t s = 1; s < n; s *= 2)
		{
			const uint32_t r_s = powmod_plain(g, (p - 1) / (6 * (uint64_t)s), p);
			if (s == 1) r = r_s;
			for (uint32_t j = 0; j < s; ++j)
			{
				const uint32_t e = 3 * bitrev_n(j, s) + 1 + (2 * j) / s;
				const uint32_t w = powmod_plain(r_s, e, p);
				const uint32_t wi = powmod_plain(w, p - 2, p);
				fw[s + j] = make_entry(to_mont(w, p), P);
				iw[s + j] = make_entry(to_mont(wi, p), P);
			}
		}
		// merge constants: inverse of the level-1 split scaled by the
		// accumulated 2^(log2 n - 2) of the unnormalised inverse butterflies and
		// by the Montgomery factor 2^-32 left by the pointwise products:
		//   u1 = (a - b) / (2r - 1), u0 = a - r u1.  We apply
		//   x[i + n/2] = (b - a) * K, x[i] = (a - (a + b) r) * K,
		//   K = 2 (1 - 2r)^-1 * (2 / n) * 2^32, stored as K * 2^32 mod p.
		const uint32_t one_m_2r = (uint32_t)(((uint64_t)p + 1 - (2 * (uint64_t)r) % p) % p);
		const uint32_t inv = powmod_plain(one_m_2r, p - 2, p);
		const uint32_t two_over_n = powmod_plain(n / 2, p - 2, p);
		uint32_t K = mulmod_plain(mulmod_plain(2, inv, p), two_over_n, p);
		K = to_mont(K, p);                 // * 2^32 (cancels the products' 2^-32)
		iw[0] = make_entry(to_mont(r, p), P);
		iw[1] = make_entry(to_mont(K, p), P);   // stored in Montgomery form
		fw[0] = make_entry(0, P);
	}
	return tw;
}

// Garner constants for the engine primes, computed on the host.
inline garner_consts make_garner()
{
	garner_consts g{};
	const uint32_t p1 = kP[0], p2 = kP[1], p3 = kP[2];
	const uint32_t i31 = powmod_plain(p3 % p1, p1 - 2, p1);
	const uint32_t i32 = powmod_plain(p3 % p2, p2 - 2, p2);
	const uint32_t i21 = powmod_plain(p2 % p1, p1 - 2, p1);
	g.inv3_1 = to_mont(i31, p1); g.inv3_1q = g.inv3_1 * kQ[0];
	g.inv3_2 = to_mont(i32, p2); g.inv3_2q = g.inv3_2 * kQ[1];
	g.inv2_1 = to_mont(i21, p1); g.inv2_1q = g.inv2_1 * kQ[0];
	g.p3_mod_p1 = p3 % p1; g.p3_mod_p2 = p3 % p2;
	const unsigned __int128 M = (unsigned __int128)p1 * p2 * p3, H = M / 2;
	g.m_lo = (uint32_t)M; g.m_mid = (uint32_t)(M >> 32); g.m_hi = (uint32_t)(M >> 64);
	g.h_lo = (uint32_t)H; g.h_mid = (uint32_t)(H >> 32); g.h_hi = (uint32_t)(H >> 64);
	g.m2 = (uint64_t)p1 * p2; g.h2 = g.m2 / 2;
	return g;
}

}   // namespace ct
