// Host-side root tables for the transform modulo X^n - X^(n/2) + 1.
//
// The factorisation tree follows cyclo/src/transform.h:336-366: level 1 splits
// X^n - X^(n/2) + 1 = (X^(n/2) - r)(X^(n/2) - r') with r a primitive 6th root of
// unity (r + r' = 1); below it, block j of level s (s blocks) is reduced by its
// root w[s + j] = r_s^(3 bitrev_s(j) + 1 + floor(2j/s)), r_s = g^((p-1)/(6s)).
// The inverse table holds w[s + j]^-1.  Entries 0/1 of the inverse table carry
// the level-1 merge constants.
//
// Table layout: tw[((P * 2) + dir) * n + idx], P = prime, dir 0 = forward,
// 1 = inverse; each entry is a Shoup pair (a, w') with w' = floor(w 2^32 / p)
// and a = 2^32 - w in the forward table (negated product, see shoup_neg), a = w
// in the inverse table.
#pragma once

#include <cstdint>
#include <thread>
#include <vector>

#include "ct_arith.h"

namespace ct {

struct tw_entry { uint32_t w, wp; };

// floor(w 2^32 / p) for w < p < 2^30, from a double estimate corrected exactly.
inline uint32_t shoup_quot(uint32_t w, uint32_t p)
{
	const unsigned __int128 num = (unsigned __int128)w << 32;
	uint64_t q = (uint64_t)((double)w * (4294967296.0 / (double)p));
	while ((unsigned __int128)q * p > num) --q;
	while ((unsigned __int128)(q + 1) * p <= num) ++q;
	return (uint32_t)q;
}

inline tw_entry fwd_entry(uint32_t w, uint32_t p) { tw_entry e; e.w = 0u - w; e.wp = shoup_quot(w, p); return e; }
inline tw_entry inv_entry(uint32_t w, uint32_t p) { tw_entry e; e.w = w; e.wp = shoup_quot(w, p); return e; }

// Tables of prime P into tw (2 * n entries at tw[P * 2 * n]).
inline void build_tables_prime(uint32_t n, int P, tw_entry *tw)
{
	const uint32_t p = kP[P], g = kG[P];
	tw_entry *fw = &tw[(size_t)(P * 2 + 0) * n];
	tw_entry *iw = &tw[(size_t)(P * 2 + 1) * n];
	uint32_t r = 0;
	std::vector<uint32_t> pw;
	// bitrev over log2(n/2) bits; for j < s, bitrev_s(j) = rev[j] >> (log2(n/2) - log2 s)
	const uint32_t half = n / 2;
	std::vector<uint32_t> rev(half);
	rev[0] = 0;
	for (uint32_t i = 1; i < half; ++i) rev[i] = (rev[i >> 1] >> 1) | ((i & 1) ? half >> 1 : 0);
	int lhalf = 0;
	while ((1u << lhalf) < half) ++lhalf;
	int ls = 0;
	for (uint32_t s = 1; s < n; s *= 2, ++ls)
	{
		// r_s has order 6s: all exponents needed are below 6s, and the
		// inverse of r_s^e is r_s^(6s - e)
		const uint32_t r_s = powmod_plain(g, (p - 1) / (6 * (uint64_t)s), p);
		if (s == 1) r = r_s;
		pw.resize(6 * (size_t)s);                 // r_s^t, t < 6s
		const uint32_t rp = shoup_quot(r_s, p);
		pw[0] = 1;
		for (size_t t = 1; t < pw.size(); ++t) pw[t] = red1p(shoup(pw[t - 1], r_s, rp, p), p);
		for (uint32_t j = 0; j < s; ++j)
		{
			const uint32_t e = 3 * (rev[j] >> (lhalf - ls)) + 1 + (2 * j) / s;
			fw[s + j] = fwd_entry(pw[e], p);
			iw[s + j] = inv_entry(pw[(6 * s - e) % (6 * s)], p);
		}
	}
	// merge constants: inverse of the level-1 split scaled by the
	// accumulated 2^(log2 n - 2) of the unnormalised inverse butterflies and
	// by the Montgomery factor 2^-32 left by the pointwise products:
	//   u1 = (a - b) / (2r - 1), u0 = a - r u1.  We apply
	//   x[i + n/2] = (b - a) * K, x[i] = (a - (a + b) r) * K,
	//   K = 2 (1 - 2r)^-1 * (2 / n) * 2^32 mod p.
	const uint32_t one_m_2r = (uint32_t)(((uint64_t)p + 1 - (2 * (uint64_t)r) % p) % p);
	const uint32_t inv = powmod_plain(one_m_2r, p - 2, p);
	const uint32_t two_over_n = powmod_plain(n / 2, p - 2, p);
	uint32_t K = mulmod_plain(mulmod_plain(2, inv, p), two_over_n, p);
	K = to_mont(K, p);                 // * 2^32 (cancels the products' 2^-32)
	iw[0] = inv_entry(r, p);
	iw[1] = inv_entry(K, p);
	fw[0].w = 0; fw[0].wp = 0;
}

// Returns 3 * 2 * n entries; one host thread per prime for large n.
inline std::vector<tw_entry> build_tables(uint32_t n)
{
	std::vector<tw_entry> tw(6 * (size_t)n);
	if (n < (1u << 16))
	{
		for (int P = 0; P < kPrimes; ++P) build_tables_prime(n, P, tw.data());
		return tw;
	}
	std::vector<std::thread> th;
	for (int P = 0; P < kPrimes; ++P) th.emplace_back(build_tables_prime, n, P, tw.data());
	for (auto &t : th) t.join();
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
